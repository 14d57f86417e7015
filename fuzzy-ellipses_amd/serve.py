"""Continuous-batching render server.

Scenes' Gaussians stay resident in HBM.  Requests (render, or render + gradient with a
given cotangent) arrive over time; the loop admits queued requests into a batch while the
HBM budget (resident scenes + in-flight workspaces, outputs and gradient buffers) allows,
launches the batch on a free stream slot and goes back to admitting while it runs, so
requests that arrive while others are on the GPU join the next batch and overlap with the
running one.  Each request in a batch is its own (scene, view) job, so its outputs and
gradients are its own; the same request alone or in a mix gives the same result.

Each request renders with the tile size tile_for(N, H, W) picks for it (unless the server
forces one), and a batch launches once per (tile size, forward / forward+backward) group.

The plain path is the same loop with one request per batch and one slot.
"""
from __future__ import annotations

import collections
import dataclasses
import math
import time
from typing import Callable, Deque, Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import Renderer, SceneTensors
from .scenes import Camera, Scene

GB = 1 << 30


@dataclasses.dataclass
class Request:
    rid: int
    scene: int
    camera: Camera
    backward: bool
    arrival: float                       # seconds after the start of the run
    d_depth: Optional[torch.Tensor] = None
    d_alpha: Optional[torch.Tensor] = None
    beta2: float = 21.4
    beta3: float = 2.66
    # filled in by the server
    depth: Optional[torch.Tensor] = None
    alpha: Optional[torch.Tensor] = None
    grads: Optional[Tuple[torch.Tensor, torch.Tensor, torch.Tensor]] = None
    t_launch: float = math.nan
    t_done: float = math.nan

    @property
    def latency_ms(self) -> float:
        return 1e3 * (self.t_done - self.arrival)


class Scheduler:
    """Continuous-batching loop, independent of the device.

    `launch(slot, batch)` starts a batch on slot `slot` and returns a handle;
    `ready(handle)` says whether it finished; `bytes_of(request)` is the memory a request
    holds while in flight.  Requests are admitted FIFO into a batch for the first idle slot
    while the free budget allows (at most `max_batch` per batch); a request that cannot fit
    waits for in-flight batches to retire.  `clock()` returns seconds; `sleep(dt)` waits.
    """

    def __init__(self, n_slots: int, max_batch: int, free_bytes: int, launch, ready, bytes_of,
                 clock: Callable[[], float] = time.perf_counter, sleep: Callable[[float], None] = time.sleep,
                 held: Optional[Callable[[], int]] = None, retire: Optional[Callable[[int], None]] = None):
        self.n_slots = max(1, int(n_slots))
        self.max_batch = max(1, int(max_batch))
        self.free_bytes = int(free_bytes)
        self.launch, self.ready, self.bytes_of = launch, ready, bytes_of
        self.clock, self.sleep = clock, sleep
        # memory held outside the in-flight requests (e.g. idle slots' workspaces), and a
        # callback told the slot whose batch retired
        self.held = held if held is not None else (lambda: 0)
        self.retire = retire if retire is not None else (lambda k: None)
        self.peak_in_flight = 0
        self.batch_sizes: List[int] = []

    def run(self, requests: Sequence[Request]) -> List[Request]:
        reqs = sorted(requests, key=lambda r: r.arrival)
        queue: Deque[Request] = collections.deque()
        slots: List[Optional[Tuple[object, List[Request], int]]] = [None] * self.n_slots
        done: List[Request] = []
        in_flight = 0
        i = 0
        t0 = self.clock()
        while len(done) < len(reqs):
            now = self.clock() - t0
            while i < len(reqs) and reqs[i].arrival <= now:
                queue.append(reqs[i])
                i += 1
            progressed = False
            for k, sl in enumerate(slots):
                if sl is not None and self.ready(sl[0]):
                    t = self.clock() - t0
                    for r in sl[1]:
                        r.t_done = t
                    done += sl[1]
                    in_flight -= sl[2]
                    slots[k] = None
                    self.retire(k)
                    progressed = True
            idle = [k for k, sl in enumerate(slots) if sl is None]
            if queue and idle:
                batch: List[Request] = []
                nbytes = 0
                held = self.held()
                while queue and len(batch) < self.max_batch:
                    b = self.bytes_of(queue[0])
                    if in_flight + held + nbytes + b > self.free_bytes:
                        break
                    batch.append(queue.popleft())
                    nbytes += b
                if batch:
                    k = idle[0]
                    t = self.clock() - t0
                    for r in batch:
                        r.t_launch = t
                    slots[k] = (self.launch(k, batch), batch, nbytes)
                    in_flight += nbytes
                    self.peak_in_flight = max(self.peak_in_flight, in_flight + self.held())
                    self.batch_sizes.append(len(batch))
                    progressed = True
                elif len(idle) == self.n_slots:
                    raise MemoryError(f"request {queue[0].rid} needs more than the free memory budget")
            if not progressed:
                busy = any(sl is not None for sl in slots)
                if busy:
                    self.sleep(2e-5)
                elif i < len(reqs):
                    self.sleep(max(0.0, min(1e-3, reqs[i].arrival - (self.clock() - t0))))
        return done


def tile_for(n: int, height: int, width: int, backward: bool = False) -> int:
    """Tile size of a request with no tile forced: 8-pixel tiles for forward-only views of up to
    320x240 pixels of scenes of 1024 Gaussians or more, 16 otherwise.  A small view's 16-pixel
    tiles span wide angles, so their lists are long and uneven and a lone request waits for its
    longest tile (one 128x128 view, N=4096: forward 0.62 -> 0.36 ms with 8-pixel tiles; 240x320:
    0.53 -> 0.35 ms; at 480x640, or with N=256, 16-pixel tiles are as fast or faster).  Requests
    with a backward keep 16-pixel tiles: the 8-pixel backward is slow in large batches (the
    saturated request mix, 2000 queued: 26.5k req/s with 16-pixel tiles throughout, 24.9k with
    8-pixel forward-only requests, 22.6k with 8-pixel backward requests too; served p50 at 800
    req/s 0.56, 0.49 and 0.49 ms).  The choice depends on the request alone, so a request gives
    the same images alone or in a mix."""
    return 8 if (not backward and height * width <= 320 * 240 and n >= 1024) else 16


class _Slot:
    """One in-flight batch position: a stream and a renderer per (tile size, fwd-only / fwd+bwd)."""

    def __init__(self, device, tiles, batch):
        self.stream = torch.cuda.Stream(device=device)
        self.renderers = {(t, bwd): Renderer(tile=t, batch=batch, mode="culled", device=device)
                          for t in tiles for bwd in (False, True)}


class Server:
    """Resident scenes + continuous batching on `slots` streams under an HBM budget."""

    def __init__(self, scenes: Sequence[Scene], device="cuda", budget_bytes: int = 64 * GB,
                 max_batch: int = 64, slots: int = 2, tile: Optional[int] = None, batch: int = 64,
                 clock: Callable[[], float] = time.perf_counter):
        """tile: the tile size of every request, or None for tile_for(N, H, W) per request."""
        self.device = torch.device(device)
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        self.scenes = [SceneTensors.from_scene(s, self.device) for s in scenes]
        self.scene_bytes = sum(s.n * 52 for s in self.scenes)
        self.budget = int(budget_bytes)
        if self.scene_bytes > self.budget:
            raise ValueError("resident scenes exceed the memory budget")
        self.max_batch = int(max_batch)
        self.tile = tile
        tiles = (8, 16) if tile is None else (int(tile),)
        self.slots = [_Slot(self.device, tiles, batch) for _ in range(max(1, slots))]
        self._slot_ws = [0] * len(self.slots)    # workspace shares of the batch in flight per slot
        self._ws_cache: Dict[tuple, int] = {}
        self.clock = clock

    def tile_of(self, req: Request) -> int:
        """The request's tile size (the forced one, else tile_for)."""
        if self.tile is not None:
            return int(self.tile)
        return tile_for(self.scenes[req.scene].n, req.camera.H, req.camera.W, bool(req.backward))

    def workspace_share(self, req: Request) -> int:
        """The workspace the planner (fe_plan) sizes for this request's view alone (with the
        backward's planes for a forward+backward request), cached per (N, H, W, tile, backward)."""
        t = self.tile_of(req)
        key = (self.scenes[req.scene].n, req.camera.H, req.camera.W, t, bool(req.backward))
        b = self._ws_cache.get(key)
        if b is None:
            rnd = self.slots[0].renderers[(t, bool(req.backward))]
            b = rnd.workspace_bytes([self.scenes[req.scene]], [(0, req.camera, req.beta2, req.beta3)],
                                    save_for_backward=req.backward)
            self._ws_cache[key] = b
        return b

    def request_bytes(self, req: Request) -> int:
        """Bytes a request holds while in flight: its workspace share (as planned by fe_plan:
        records, masks, worst-case tile lists, per-pixel state, backward planes), outputs and,
        for a backward, its gradient buffers (the cotangents are the caller's)."""
        n = self.scenes[req.scene].n
        b = self.workspace_share(req) + 8 * req.camera.H * req.camera.W       # + depth, alpha
        if req.backward:
            b += 52 * n                                                          # d_means, d_prec, d_w
        return b + 4096

    def held_bytes(self) -> int:
        """Workspace the slots' renderers keep (at their high-water mark) beyond the shares of
        the requests now in flight on them."""
        tot = 0
        for k, slot in enumerate(self.slots):
            cap = sum(r._ws.numel() if r._ws is not None else 0 for r in slot.renderers.values())
            tot += max(0, cap - self._slot_ws[k])
        return tot

    def _launch(self, k: int, batch: List[Request]):
        slot = self.slots[k]
        self._slot_ws[k] = sum(self.workspace_share(r) for r in batch)
        with torch.cuda.stream(slot.stream):
            groups: Dict[tuple, List[Request]] = {}
            for r in batch:
                groups.setdefault((self.tile_of(r), bool(r.backward)), []).append(r)
            for (t, bwd), g in sorted(groups.items()):
                rnd = slot.renderers[(t, bwd)]
                scenes = [self.scenes[r.scene] for r in g]
                views = [(i, r.camera, r.beta2, r.beta3) for i, r in enumerate(g)]
                outs = rnd.alloc_outputs(views)
                rnd.forward(scenes, views, save_for_backward=bwd, outs=outs)
                for r, (d, a) in zip(g, outs):
                    r.depth, r.alpha = d, a
                if bwd:
                    grads = rnd.backward([r.d_depth for r in g], [r.d_alpha for r in g])
                    for r, gr in zip(g, grads):
                        r.grads = gr
            ev = torch.cuda.Event()
            ev.record(slot.stream)
        return ev

    def _retire(self, k: int) -> None:
        self._slot_ws[k] = 0

    def run(self, requests: Sequence[Request]) -> Dict[str, float]:
        torch.cuda.synchronize(self.device)      # inputs (cotangents) are ready
        sch = Scheduler(len(self.slots), self.max_batch, self.budget - self.scene_bytes, self._launch,
                        lambda ev: ev.query(), self.request_bytes, clock=self.clock, held=self.held_bytes,
                        retire=self._retire)
        t0 = self.clock()
        done = sch.run(requests)
        total = self.clock() - t0
        lat = np.array([r.latency_ms for r in done])
        pix = sum(r.camera.H * r.camera.W for r in done)
        return {"requests": len(done), "batches": len(sch.batch_sizes), "seconds": total,
                "mean_batch": float(np.mean(sch.batch_sizes)) if sch.batch_sizes else 0.0,
                "throughput_rps": len(done) / total, "megapixels_per_s": pix / total / 1e6,
                "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                "mean_ms": float(lat.mean()), "peak_in_flight_bytes": sch.peak_in_flight}


# ------------------------------------------------------------------------------------------
# synthetic request streams
# ------------------------------------------------------------------------------------------
def poisson_requests(scenes: Sequence[Scene], rate: float, n: int, seed: int = 4, backward_frac: float = 0.3,
                     resolutions=((128, 128), (240, 320), (480, 640)), device="cuda",
                     burst: bool = False) -> List[Request]:
    """Poisson arrivals at `rate` req/s (all at t=0 if `burst`), scene uniform, resolution
    uniform over `resolutions`, camera on a random direction at radius U[2.5, 3.5] looking
    at the origin (f = W * 600/640), a `backward_frac` share asking for gradients with a
    random cotangent (one shared cotangent per resolution)."""
    from .scenes import make_camera, random_dir
    rng = np.random.default_rng(seed)
    dev = torch.device(device)
    cot = {}
    g = torch.Generator(device="cpu").manual_seed(seed)
    for (H, W) in resolutions:
        cot[(H, W)] = (torch.randn(H, W, generator=g).to(dev), torch.randn(H, W, generator=g).to(dev))
    t = 0.0
    out = []
    for k in range(n):
        if not burst:
            t += float(rng.exponential(1.0 / rate))
        si = int(rng.integers(len(scenes)))
        H, W = resolutions[int(rng.integers(len(resolutions)))]
        cam = make_camera(float(rng.uniform(2.5, 3.5)) * random_dir(rng), H, W, W * 600.0 / 640.0)
        bwd = bool(rng.random() < backward_frac)
        d_d, d_a = cot[(H, W)] if bwd else (None, None)
        out.append(Request(k, si, cam, bwd, t, d_d, d_a, scenes[si].beta2, scenes[si].beta3))
    return out
