"""Batched launch, HIP-graph replay and autotuner invariance on the GPU."""
import numpy as np
import pytest

from oracle import ref

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def test_batched_equals_separate(fea, scenes_mod):
    """Many scenes of different sizes and resolutions in one call give bit-identical images
    to one call per view, and match the FP64 oracle on sampled pixels."""
    import torch
    scenes = scenes_mod.batched_scenes(n_scenes=6)
    sts = [fea.SceneTensors.from_scene(s) for s in scenes]
    views = [(i, c, s.beta2, s.beta3) for i, s in enumerate(scenes) for c in s.cameras]
    r = fea.Renderer(tile=16, batch=64)
    outs = r.forward(sts, views)
    torch.cuda.synchronize()
    rng = np.random.default_rng(5)
    for k, (si, cam, b2, b3) in enumerate(views):
        r1 = fea.Renderer(tile=16, batch=64)
        (d1, a1), = r1.forward([sts[si]], [(0, cam, b2, b3)])
        torch.cuda.synchronize()
        np.testing.assert_array_equal(_np(outs[k][0]), _np(d1))
        np.testing.assert_array_equal(_np(outs[k][1]), _np(a1))
        if k % 4 == 0:
            pix = rng.choice(cam.H * cam.W, min(400, cam.H * cam.W), replace=False)
            od, oa = ref.render_f64(scenes[si], cam, pixels=pix)
            d = _np(d1).ravel()[pix]
            a = _np(a1).ravel()[pix]
            assert np.max(np.abs(a - oa)) <= 1e-4
            assert np.max(np.abs(d - od) / np.abs(od)) <= 1e-4
    print(f"batched: {len(views)} views of {len(scenes)} scenes in one call")


def test_graph_replay_matches_eager(fea, scenes_mod):
    """Forward+backward captured in a HIP graph (descriptors uploaded beforehand) replays the
    eager result, replay after replay, with eager work in between."""
    import torch
    scene = scenes_mod.single_view_scene(seed=9, n=200, H=96, W=128, f=100.0)
    st = fea.SceneTensors.from_scene(scene)
    cams = [scene.cameras[0]] + [scenes_mod.make_camera(2.5 * d, 96, 128, 100.0)
                                 for d in scenes_mod.fibonacci_dirs(3)]
    views = [(0, c, scene.beta2, scene.beta3) for c in cams]
    r = fea.Renderer(tile=16, batch=64)
    outs = r.alloc_outputs(views)
    g = torch.Generator(device="cpu").manual_seed(1)
    dd = [torch.randn(c.H, c.W, generator=g).cuda() for c in cams]
    da = [torch.randn(c.H, c.W, generator=g).cuda() for c in cams]
    grads = [(torch.zeros_like(st.means), torch.zeros_like(st.prec), torch.zeros_like(st.weights_log))]

    def step(reuse):
        for t in grads[0]:
            t.zero_()
        r.forward([st], views, save_for_backward=True, outs=outs, reuse_descriptors=reuse)
        r.backward(dd, da, grads=grads, reuse_descriptors=reuse)

    step(False)
    torch.cuda.synchronize()
    ref_img = [(_np(d).copy(), _np(a).copy()) for d, a in outs]
    ref_g = [_np(x).copy() for x in grads[0]]
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        step(False)
    torch.cuda.current_stream().wait_stream(side)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        step(True)
    for rep in range(3):
        for d, a in outs:          # eager work between replays
            d.fill_(-1.0)
            a.fill_(-1.0)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        for (d, a), (rd, ra) in zip(outs, ref_img):
            np.testing.assert_array_equal(_np(d), rd)
            np.testing.assert_array_equal(_np(a), ra)
        for x, rx in zip(grads[0], ref_g):
            np.testing.assert_allclose(_np(x), rx, rtol=1e-4, atol=1e-6 * np.abs(rx).max())


def test_tile_and_batch_do_not_change_results(fea, scenes_mod):
    import torch
    scene = scenes_mod.shell_scene(n_views=2, H=240, W=320, f=300.0)
    st = fea.SceneTensors.from_scene(scene)
    views = [(0, c, scene.beta2, scene.beta3) for c in scene.cameras]
    res = {}
    for tile in (8, 16, 32):
        for batch in (32, 256):
            r = fea.Renderer(tile=tile, batch=batch)
            outs = r.forward([st], views)
            torch.cuda.synchronize()
            res[(tile, batch)] = [(_np(d), _np(a)) for d, a in outs]
    base = res[(16, 64)] if (16, 64) in res else res[(16, 32)]
    worst = max(float(np.max(np.abs(d - bd) / np.abs(bd))) for v in res.values() for (d, _), (bd, _) in zip(v, base))
    print(f"tile/batch configs: largest depth difference {worst:.3e} (relative)")
    for k, v in res.items():
        for (d, a), (bd, ba) in zip(v, base):
            assert np.max(np.abs(a - ba)) <= 1e-6, k
            assert np.max(np.abs(d - bd) / np.abs(bd)) <= 2e-6, k


def test_autotuner_picks_a_swept_config(fea, scenes_mod, tmp_path):
    import torch
    from fuzzy_ellipses_amd.autotune import Autotuner, workload_key
    scene = scenes_mod.shell_scene(n=1024, n_views=2, H=240, W=320, f=300.0)
    st = fea.SceneTensors.from_scene(scene)
    views = [(0, c, scene.beta2, scene.beta3) for c in scene.cameras]
    at = Autotuner(path=str(tmp_path / "tune.json"), reps=2)
    best = at.tune([st], views, backward=True, tiles=(8, 16, 32), batches=(32, 128))
    e = at.table[workload_key([st], views, True)]
    assert (best[0], best[1]) == (e["tile"], e["batch"])
    assert len(e["sweep"]) == 6 and e["ms"] == min(e["sweep"].values()) or abs(e["ms"] - min(e["sweep"].values())) < 1e-3
    at2 = Autotuner(path=str(tmp_path / "tune.json"))
    assert at2.best([st], views, backward=True) == best
    print("autotune sweep (ms):", e["sweep"])
