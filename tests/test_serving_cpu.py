"""Continuous-batching scheduler (fuzzy-ellipses_amd/serve.py) on a fake device and clock."""
import numpy as np
import pytest


class FakeDevice:
    """Batches take `per_req` seconds per request of simulated time."""

    def __init__(self, clock, per_req=0.01):
        self.clock, self.per_req = clock, per_req
        self.launched = []

    def launch(self, slot, batch):
        self.launched.append((slot, [r.rid for r in batch], self.clock.t))
        return self.clock.t + self.per_req * len(batch)

    def ready(self, handle):
        return self.clock.t >= handle


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += max(dt, 1e-4)


def _reqs(fea, n, rate=None, bwd_every=3):
    from fuzzy_ellipses_amd.serve import Request
    import fuzzy_ellipses_amd.scenes as sc
    cam = sc.make_camera([0, 0, 3.0], 8, 8, 8.0)
    rng = np.random.default_rng(0)
    t = 0.0
    out = []
    for k in range(n):
        if rate:
            t += rng.exponential(1 / rate)
        out.append(Request(k, k % 3, cam, k % bwd_every == 0, t))
    return out


def _sched(fea, clock, dev, slots=2, max_batch=8, free=10 ** 9, bytes_of=lambda r: 100):
    from fuzzy_ellipses_amd.serve import Scheduler
    return Scheduler(slots, max_batch, free, dev.launch, dev.ready, bytes_of, clock=clock, sleep=clock.sleep)


def test_every_request_completes_and_batches_form_under_load(fea):
    clock = FakeClock()
    dev = FakeDevice(clock, per_req=0.002)
    reqs = _reqs(fea, 200, rate=2000)
    done = _sched(fea, clock, dev).run(reqs)
    assert sorted(r.rid for r in done) == list(range(200))
    assert all(r.t_done >= r.t_launch >= r.arrival for r in done)
    sizes = [len(b) for _, b, _ in dev.launched]
    assert max(sizes) > 1          # arrivals during a running batch were batched together
    assert max(sizes) <= 8


def test_continuous_admission_overlaps_slots(fea):
    clock = FakeClock()
    dev = FakeDevice(clock, per_req=0.01)
    reqs = _reqs(fea, 40, rate=400)
    _sched(fea, clock, dev, slots=2).run(reqs)
    # some batch was launched while the other slot was still busy
    starts = [(s, t) for s, _, t in dev.launched]
    assert {s for s, _ in starts} == {0, 1}


def test_memory_budget_limits_batches(fea):
    clock = FakeClock()
    dev = FakeDevice(clock)
    reqs = _reqs(fea, 30)
    sch = _sched(fea, clock, dev, slots=2, max_batch=64, free=350, bytes_of=lambda r: 100)
    done = sch.run(reqs)
    assert len(done) == 30
    assert sch.peak_in_flight <= 350
    assert max(sch.batch_sizes) <= 3


def test_request_larger_than_budget_is_rejected(fea):
    clock = FakeClock()
    dev = FakeDevice(clock)
    with pytest.raises(MemoryError):
        _sched(fea, clock, dev, free=50, bytes_of=lambda r: 100).run(_reqs(fea, 2))


def test_plain_path_is_one_request_per_launch(fea):
    clock = FakeClock()
    dev = FakeDevice(clock)
    done = _sched(fea, clock, dev, slots=1, max_batch=1).run(_reqs(fea, 10, rate=1000))
    assert len(done) == 10 and all(len(b) == 1 for _, b, _ in dev.launched)


def test_held_memory_counts_against_the_budget(fea):
    """Memory held outside the in-flight requests (idle slots' workspaces) shrinks what a batch
    may take; the retire callback names each slot whose batch finished."""
    clock = FakeClock()
    dev = FakeDevice(clock)
    from fuzzy_ellipses_amd.serve import Scheduler
    retired = []
    sch = Scheduler(2, 64, 600, dev.launch, dev.ready, lambda r: 100, clock=clock, sleep=clock.sleep,
                    held=lambda: 300, retire=retired.append)
    done = sch.run(_reqs(fea, 12))
    assert len(done) == 12
    assert max(sch.batch_sizes) <= 3 and sch.peak_in_flight <= 600
    assert len(retired) == len(sch.batch_sizes) and set(retired) <= {0, 1}


def test_planner_sizes_what_requests_are_charged(fea):
    """Server.request_bytes charges fe_plan's workspace for the request's view (planning is host
    arithmetic: no GPU needed); that workspace holds at least the worst-case tile lists, 4 bytes
    per (tile, Gaussian), with or without the backward's planes."""
    import ctypes
    from fuzzy_ellipses_amd import _lib
    import fuzzy_ellipses_amd.scenes as sc
    lib = _lib.load()
    cam = sc.make_camera([0, 0, 3.0], 240, 320, 300.0)
    for n, bwd in ((500, False), (4096, True)):
        scn = (_lib.FeScene * 1)(_lib.FeScene(n, None, None, None))
        c = _lib.FeCamera()
        R = np.asarray(cam.R, np.float32).ravel()
        for i in range(9):
            c.R[i] = float(R[i])
        c.fx, c.fy, c.cx, c.cy, c.height, c.width = cam.fx, cam.fy, cam.cx, cam.cy, cam.H, cam.W
        vw = (_lib.FeView * 1)(_lib.FeView(0, c, 21.4, 2.66, None, None))
        nbytes = ctypes.c_size_t(0)
        assert lib.fe_plan(scn, 1, vw, 1, ctypes.byref(_lib.FeOptions(16, 64, 0, int(bwd), 0)),
                           ctypes.byref(nbytes)) == 0
        # the planner's worst-case lists alone are 4 bytes per (tile, Gaussian)
        assert nbytes.value >= 4 * n * 15 * 20


def test_tile_policy_depends_on_the_request_alone(fea):
    """Requests with no tile forced: 8-pixel tiles for forward-only views of up to 320x240
    pixels of scenes of 1024 Gaussians or more, 16-pixel tiles otherwise."""
    from fuzzy_ellipses_amd.serve import tile_for
    assert tile_for(4096, 128, 128) == 8 and tile_for(1024, 240, 320) == 8
    assert tile_for(4096, 480, 640) == 16 and tile_for(1023, 128, 128) == 16 and tile_for(4096, 241, 320) == 16
    assert tile_for(4096, 128, 128, backward=True) == 16
