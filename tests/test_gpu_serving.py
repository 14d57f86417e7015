"""Serving on the GPU: each request's result is the same alone and in a continuous-batching
mix (images bit-identical; gradients within 1e-5 relative, float32 atomics reorder sums)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def test_alone_equals_in_mix(fea, scenes_mod):
    import torch
    from fuzzy_ellipses_amd.serve import Server, poisson_requests
    # N up to 2048: the batches mix 8-pixel-tile requests (N >= 1024) with 16-pixel ones
    scenes = scenes_mod.resident_scenes(n_scenes=4, n_lo=64, n_hi=2048)
    reqs = poisson_requests(scenes, rate=0, n=24, seed=11, backward_frac=0.4,
                            resolutions=((64, 64), (96, 128)), burst=True)
    mix = Server(scenes, max_batch=8, slots=2)
    assert {mix.tile_of(r) for r in reqs} == {8, 16}
    st = mix.run(reqs)
    assert st["requests"] == 24 and st["mean_batch"] > 1
    mixed = {r.rid: (_np(r.depth).copy(), _np(r.alpha).copy(),
                     [_np(g).copy() for g in r.grads] if r.backward else None) for r in reqs}
    alone = Server(scenes, max_batch=1, slots=1)
    reqs2 = poisson_requests(scenes, rate=0, n=24, seed=11, backward_frac=0.4,
                             resolutions=((64, 64), (96, 128)), burst=True)
    alone.run(reqs2)
    for r in reqs2:
        d, a, g = mixed[r.rid]
        np.testing.assert_array_equal(_np(r.depth), d)
        np.testing.assert_array_equal(_np(r.alpha), a)
        if r.backward:
            for x, y in zip(r.grads, g):
                x = _np(x)
                assert np.max(np.abs(x - y)) <= 1e-5 * max(np.max(np.abs(y)), 1e-30)


def test_poisson_serving_reports_latency(fea, scenes_mod):
    from fuzzy_ellipses_amd.serve import Server, poisson_requests
    scenes = scenes_mod.resident_scenes(n_scenes=8, n_lo=64, n_hi=1024)
    reqs = poisson_requests(scenes, rate=400, n=200, seed=12, resolutions=((128, 128), (240, 320)))
    st = Server(scenes, max_batch=32, slots=2).run(reqs)
    print("serving:", {k: round(v, 3) if isinstance(v, float) else v for k, v in st.items()})
    assert st["requests"] == 200
    assert 0 < st["p50_ms"] <= st["p99_ms"]
    assert all(r.depth is not None for r in reqs)
