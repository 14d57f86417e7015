"""Benchmark of the fuzzy-metaball rasterizer hot path on one MI355X.

Workload (the headline configuration): N=4096 Gaussians on a unit-sphere shell, 16
fibonacci views at radius 3, 480x640 pixels, f=600, beta2=21.4, beta3=2.66, float32.  One
step = one batched forward of all 16 views (saving aggregates) + the hand-written backward
of a fixed random cotangent for depth and alpha (the shape-from-silhouette step).  Inputs
are resident in HBM before the timed region.

Prints one JSON line.  `value` = forward+backward frames/s (16 frames per step).  The line
also carries the forward-only rates, ray-Gaussian evaluation rates, the roofline of the
dominant kernel (per-kernel HIP-event timing on the launch stream over the timed region) and
a CPU baseline timed on a bounded sample.
"""
import argparse
import importlib.util
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "fuzzy-ellipses_amd")

METRIC = "ray-Gaussian evals/s + fwd+bwd frames/s @480x640,N=4096; served p50/p99 ms"
FP32_PEAK_TFLOPS = 157.3          # MI355X published dense FP32 vector peak
FLOPS_FWD = 36                    # per evaluated (pixel, Gaussian) pair, blend forward
FLOPS_BWD = 93                    # per evaluated pair, blend backward incl. recompute
TRANS_PER_PAIR = 3                # v_exp x2 + v_rcp per evaluated pair, forward and backward
TRANS_PEAK_TOPS = 9.8             # 256 CUs x 4 SIMDs x 16 lanes / 4 (quarter rate) x 2.4 GHz
HBM_PEAK_GBS = 8000.0             # MI355X published HBM3E bandwidth
TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                            "traffic_shell_n4096_16x480x640.json")
# kernel names in the profile (the 16-view workload's 480x640 views split their phase-1 lists,
# one range after the other on a grid this size: mode 1; the second phase never splits)
PROFILE_NAMES = {"bwd_blend": "fe::k_bwd_blend<16>", "blend1": "fe::k_blend<16, 64, 1, 1>",
                 "blend2": "fe::k_blend<16, 64, 2, 0>"}


def profiled_counters(dom, args):
    """Per-dispatch counters of kernel `dom` on this workload from the committed rocprofv3 --pmc
    summary (HBM bytes: FETCH_SIZE x2 + WRITE_SIZE x1, calibrated on a copy kernel; VALU wave
    instructions), or None for a configuration it does not cover."""
    if args.tile != 16 or args.batch != 64 or args.mode != "culled" or not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as f:
        return json.load(f)["kernels"].get(PROFILE_NAMES.get(dom, ""))


def roofline(dom, avg_s, pairs, flops_per_pair, alg_bytes, args):
    """Roofline record of one kernel: algorithmic FP32 flops (the hot-path spec's per-pair figure)
    and HBM bytes against MI355X's peaks, with the issued-instruction view beside them."""
    achieved = flops_per_pair * pairs / avg_s / 1e12 if avg_s > 0 else 0.0
    trans = TRANS_PER_PAIR * pairs / avg_s / 1e12 if avg_s > 0 else 0.0
    prof = profiled_counters(dom, args)
    traffic = None if prof is None else int(prof["read_bytes"] + prof["write_bytes"])
    out = {"bound": "fp32_valu", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
           "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": traffic, "kernel": dom,
           "launch_avg_ms": round(avg_s * 1e3, 5), "pairs_per_launch": int(pairs),
           "flops_per_pair": flops_per_pair,
           "flops_model": "algorithmic FP32 flops per evaluated (pixel, Gaussian) pair, FMA = 2 (no MFMA: "
                          "per-ray 3x3 forms are not a dense contraction)",
           "trans_achieved": round(trans, 3), "trans_peak": TRANS_PEAK_TOPS, "trans_unit": "Top/s",
           "trans_frac": round(trans / TRANS_PEAK_TOPS, 4),
           "algorithmic_bytes": int(alg_bytes),
           "hbm_algorithmic_gbs": round(alg_bytes / avg_s / 1e9, 2) if avg_s > 0 else 0.0,
           "hbm_peak_gbs": HBM_PEAK_GBS}
    if traffic is not None:
        out.update(traffic_source=os.path.relpath(TRAFFIC_FILE, ROOT) + " (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE "
                   "per dispatch of this workload, copy-calibrated)",
                   hbm_achieved_gbs=round(traffic / avg_s / 1e9, 2),
                   hbm_frac=round(traffic / avg_s / 1e9 / HBM_PEAK_GBS, 4),
                   traffic_over_algorithmic=round(traffic / alg_bytes, 3))
    if prof is not None and "valu_insts" in prof:
        vpp = prof["valu_insts"] * 64 / pairs
        issued = 2.0 * prof["valu_insts"] * 64 / avg_s / 1e12
        out.update(issued_valu_per_pair=round(vpp, 2),
                   issued_flops_upper_tflops=round(issued, 3),
                   issued_frac_upper=round(issued / FP32_PEAK_TFLOPS, 4),
                   issued_note="SQ_INSTS_VALU x 64 lanes per evaluated pair from the same profile; the upper "
                               "bound counts every issued VALU instruction as an FMA (2 flops per lane)")
    return out


def _package():
    if "fuzzy_ellipses_amd" in sys.modules:
        return sys.modules["fuzzy_ellipses_amd"]
    spec = importlib.util.spec_from_file_location(
        "fuzzy_ellipses_amd", os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["fuzzy_ellipses_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def cpu_baseline(scene, cam, budget_s=12.0):
    """fwd+bwd of the float64 autograd restatement on a bounded sample of pixels of one view,
    on the host cores; scaled to frames/s of the full 480x640 view."""
    import numpy as np
    import torch
    sys.path.insert(0, ROOT)
    from oracle import grad_oracle as go  # the CPU reference restatement

    cores = min(os.cpu_count() or 1, 16)
    torch.set_num_threads(cores)
    rng = np.random.default_rng(0)
    means = torch.tensor(scene.means, requires_grad=True)
    prec = torch.tensor(scene.prec, requires_grad=True)
    w = torch.tensor(scene.weights_log, requires_grad=True)
    chunk = 512
    done = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s and done < cam.H * cam.W:
        pix = rng.choice(cam.H * cam.W, chunk, replace=False)
        rays = torch.tensor(go.pinhole_rays(cam, pix))
        starts = torch.tensor(np.repeat(np.asarray(cam.t, dtype=np.float64)[None], chunk, 0))
        d, a = go.render_rays_torch(rays, starts, means, prec, w, scene.beta2, scene.beta3)
        (d.sum() + a.sum()).backward()
        done += chunk
    el = time.perf_counter() - t0
    return {"value": (done / (cam.H * cam.W)) / el, "unit": "frames/s", "cores": cores, "kind": "port",
            "sample": f"{done} random pixels of one 480x640 view, N=4096, float64 torch autograd fwd+bwd "
                      f"of the restated render_func_rays, {el:.1f} s"}


def cpu_forward_f32(scene, cam, rows=32):
    """Forward-only float32 C++ (OpenMP) dense port on `rows` image rows -> frames/s."""
    import numpy as np
    sys.path.insert(0, ROOT)
    from oracle import ref
    cores = min(os.cpu_count() or 1, 16)
    pix = np.arange(rows * cam.W, dtype=np.int64) + (cam.H // 2 - rows // 2) * cam.W
    ref.render_f32(scene, cam, pixels=pix[: cam.W], nthreads=cores)
    t0 = time.perf_counter()
    ref.render_f32(scene, cam, pixels=pix, nthreads=cores)
    el = time.perf_counter() - t0
    return {"value": (len(pix) / (cam.H * cam.W)) / el, "unit": "frames/s", "cores": cores, "kind": "port",
            "sample": f"{rows} rows of one view, dense float32 forward, {el:.2f} s"}


def served(fea, sc, dev, rate=800.0, n=1000, burst=False, max_batch=64, slots=2, seed=4,
           resolutions=((128, 128), (240, 320), (480, 640)), backward_frac=0.3, scenes=None):
    """Serving run of the request mix over 64 resident scenes -> stats dict."""
    from fuzzy_ellipses_amd.serve import Server, poisson_requests
    scenes = scenes if scenes is not None else sc.resident_scenes()
    srv = Server(scenes, device=dev, max_batch=max_batch, slots=slots)
    warm = poisson_requests(scenes, rate, min(64, n), seed=seed + 100, backward_frac=backward_frac,
                            resolutions=resolutions, device=dev, burst=True)
    srv.run(warm)
    reqs = poisson_requests(scenes, rate, n, seed=seed, backward_frac=backward_frac, resolutions=resolutions,
                            device=dev, burst=burst)
    st = srv.run(reqs)
    return {k: (round(v, 4) if isinstance(v, float) else v) for k, v in st.items()}


def batched_summary(fea, sc, dev, steps, tune_path=None):
    """256 scenes (N ~ U{16..2048}) with 1-8 views each at mixed resolutions: one call per
    kernel for the whole batch (autotuned tile and LDS batch), the same call replayed from a
    HIP graph, and one call per view (the plain path)."""
    import torch
    from fuzzy_ellipses_amd.autotune import Autotuner
    scenes = sc.batched_scenes()
    sts = [fea.SceneTensors.from_scene(s, dev) for s in scenes]
    views = [(i, c, s.beta2, s.beta3) for i, s in enumerate(scenes) for c in s.cameras]
    at = Autotuner(dev, path=tune_path, reps=2)
    tile, batch = at.tune(sts, views, backward=False, force=True)
    r = fea.Renderer(tile=tile, batch=batch, device=dev)
    outs = r.alloc_outputs(views)
    r.forward(sts, views, outs=outs)
    torch.cuda.synchronize()
    stats = r.stats()
    t0 = time.perf_counter()
    for _ in range(steps):
        r.forward(sts, views, outs=outs, reuse_descriptors=True)
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / steps
    # the same views through a captured HIP graph
    graph = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream(dev)
    side.wait_stream(torch.cuda.current_stream(dev))
    with torch.cuda.stream(side):
        r.forward(sts, views, outs=outs)
    torch.cuda.current_stream(dev).wait_stream(side)
    with torch.cuda.graph(graph):
        r.forward(sts, views, outs=outs, reuse_descriptors=True)
    graph.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        graph.replay()
    torch.cuda.synchronize()
    el_g = (time.perf_counter() - t0) / steps
    # plain path: one view per call
    r1 = fea.Renderer(tile=16, batch=64, device=dev)
    for (si, cam, b2, b3) in views[:8]:
        r1.forward([sts[si]], [(0, cam, b2, b3)])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for (si, cam, b2, b3) in views:
        r1.forward([sts[si]], [(0, cam, b2, b3)])
    torch.cuda.synchronize()
    el_1 = time.perf_counter() - t0
    dense = sum(sts[si].n * cam.H * cam.W for (si, cam, _, _) in views)
    key = next(iter(at.table)) if at.table else None
    return {"workload": "256 scenes N~U{16..2048}, 1-8 views each, resolutions 64^2..480x640, forward",
            "views": len(views), "tile": tile, "lds_batch": batch, "steps": steps,
            "frames_per_s": round(len(views) / el, 2), "ms_per_call": round(1e3 * el, 3),
            "graph_replay_frames_per_s": round(len(views) / el_g, 2),
            "plain_one_view_per_call_frames_per_s": round(len(views) / el_1, 2),
            "speedup_vs_plain": round(el_1 / el, 2),
            "evals_per_s": stats["evals"] / el, "dense_equiv_evals_per_s": dense / el,
            "autotune_sweep_ms": at.table[key]["sweep"] if key else None}


def workload_batched(fea, sc, dev, steps):
    b = batched_summary(fea, sc, dev, steps)
    return {"metric": "batched forward frames/s (256 scenes, mixed N and resolution)", "value": b["frames_per_s"],
            "unit": "frames/s", "n_gpus": 1, "steps": steps, "ms_per_step": b["ms_per_call"],
            "higher_is_better": True, "dtype": "f32", "data": "synthetic",
            "config": {"workload": b["workload"], "views": b["views"], "tile": b["tile"], "lds_batch": b["lds_batch"]},
            **{k: v for k, v in b.items() if k not in ("workload", "views", "tile", "lds_batch", "steps")}}


def serving_summary(fea, sc, dev, rates, n_per_rate):
    """Poisson request mix (70% forward, 30% forward+backward) over 64 resident scenes at each
    rate, beside the plain one-request-per-launch path at the same rate; the saturated
    throughput of the batched server vs the plain path for fwd+bwd at 480x640, N=4096 and for
    the request mix itself (every request queued at once)."""
    out = {"workload": "64 resident scenes N~U{64..4096}, resolutions 128^2/240x320/480x640, 70% fwd / "
                       "30% fwd+bwd, Poisson arrivals", "rates": {}}
    for rate, n in zip(rates, n_per_rate):
        e = served(fea, sc, dev, rate=rate, n=n)
        plain = served(fea, sc, dev, rate=rate, n=max(n // 3, 50), max_batch=1, slots=1)
        e.update(plain_requests=plain["requests"], plain_p50_ms=plain["p50_ms"], plain_p99_ms=plain["p99_ms"],
                 plain_throughput_rps=plain["throughput_rps"])
        out["rates"][str(int(rate)) if float(rate).is_integer() else str(rate)] = e
    shell = sc.shell_scene(n_views=1)
    big = [sc.Scene(shell.means, shell.prec, shell.weights_log, [])]
    sat_b = served(fea, sc, dev, rate=1.0, n=160, burst=True, max_batch=16, slots=2, resolutions=((480, 640),),
                   backward_frac=1.0, scenes=big)
    sat_p = served(fea, sc, dev, rate=1.0, n=40, burst=True, max_batch=1, slots=1, resolutions=((480, 640),),
                   backward_frac=1.0, scenes=big)
    out["saturated_fwd_bwd_480x640_n4096"] = {"requests": 160, "batched_rps": sat_b["throughput_rps"],
                                              "plain_requests": 40, "plain_rps": sat_p["throughput_rps"],
                                              "ratio": round(sat_b["throughput_rps"] / sat_p["throughput_rps"], 3)}
    mix_b = served(fea, sc, dev, rate=1.0, n=2000, burst=True, max_batch=256)
    mix_p = served(fea, sc, dev, rate=1.0, n=400, burst=True, max_batch=1, slots=1)
    out["saturated_mix"] = {"requests": 2000, "batched_rps": mix_b["throughput_rps"], "batched_p50_ms": mix_b["p50_ms"],
                            "batched_p99_ms": mix_b["p99_ms"], "mean_batch": mix_b["mean_batch"],
                            "plain_requests": 400, "plain_rps": mix_p["throughput_rps"],
                            "ratio": round(mix_b["throughput_rps"] / mix_p["throughput_rps"], 3)}
    out["target_20x_met"] = bool(out["saturated_fwd_bwd_480x640_n4096"]["ratio"] >= 20.0)
    return out


def workload_serving(fea, sc, dev, rates, n_per_rate):
    out = {"metric": "served p50/p99 ms and throughput", "unit": "ms", "n_gpus": 1, "data": "synthetic"}
    s = serving_summary(fea, sc, dev, rates, n_per_rate)
    out["config"] = {"workload": s.pop("workload")}
    out.update(s)
    out["value"] = out["rates"][list(out["rates"])[-1]]["p99_ms"]
    out["higher_is_better"] = False
    return out


GMM_FLOPS_FWD = 21      # per (point, Gaussian) pair: p (3), A p (9), |.|^2 (5), log term, online max/sum (4)
GMM_FLOPS_BWD = 38      # recompute (18), responsibility (2), moments (18)


def workload_gmm(fea, sc, dev, steps):
    """GMM point-cloud NLL, forward + backward: N=4096 Gaussians, 2^20 points."""
    import numpy as np
    import torch
    from fuzzy_ellipses_amd.gmm import GMM
    rng = np.random.default_rng(5)
    means, U, wl = sc.random_gaussians(rng, 4096)
    g = GMM(means, U, wl, device=dev)
    m = 1 << 20
    pts = torch.tensor(rng.uniform(-0.6, 0.6, size=(m, 3)), dtype=torch.float32, device=dev)
    lse = torch.empty(m, dtype=torch.float32, device=dev)
    nll = torch.empty((), dtype=torch.float32, device=dev)
    grads = (torch.zeros_like(g.means), torch.zeros_like(g.prec), torch.zeros_like(g.weights_log))
    s = torch.cuda.current_stream(dev)
    for _ in range(2):
        g.forward(pts, lse=lse, nll=nll)
        g.backward(1.0, grads)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    tf = tb = 0.0
    for _ in range(steps):
        ev[0].record(s)
        g.forward(pts, lse=lse, nll=nll)
        ev[1].record(s)
        g.backward(1.0, grads)
        ev[2].record(s)
        torch.cuda.synchronize()
        tf += ev[0].elapsed_time(ev[1])
        tb += ev[1].elapsed_time(ev[2])
    pairs = 4096 * m
    fwd_s, bwd_s = tf / steps / 1e3, tb / steps / 1e3
    return {"metric": "GMM point-cloud NLL fwd+bwd (point, Gaussian) pairs/s (N=4096, M=2^20)",
            "value": round(pairs / (fwd_s + bwd_s), 1), "unit": "pairs/s", "n_gpus": 1, "steps": steps,
            "higher_is_better": True, "dtype": "f32", "data": "synthetic",
            "fwd_ms": round(fwd_s * 1e3, 4), "bwd_ms": round(bwd_s * 1e3, 4),
            "fwd_tflops": round(GMM_FLOPS_FWD * pairs / fwd_s / 1e12, 3),
            "bwd_tflops": round(GMM_FLOPS_BWD * pairs / bwd_s / 1e12, 3),
            "peak_fp32_tflops": FP32_PEAK_TFLOPS,
            "config": {"workload": "GMM log-likelihood of 2^20 uniform points under 4096 Gaussians"}}


def workload_silhouette(fea, sc, dev, steps):
    from fuzzy_ellipses_amd.silhouette import SilhouetteFit
    init, target = sc.silhouette_problem()
    fit = SilhouetteFit(init, target, device=dev, lr=1e-2)
    fit.run(steps=2)
    res = fit.run(steps=steps)
    return {"metric": "shape-from-silhouette fwd+bwd+Adam frames/s (N=1024, 16 views 240x320)",
            "value": round(res["frames_per_s"], 2), "unit": "frames/s", "n_gpus": 1, "steps": steps,
            "ms_per_step": round(res["ms_per_step"], 3), "higher_is_better": True, "dtype": "f32",
            "data": "synthetic", "loss_first": res["loss_first"], "loss_last": res["loss_last"],
            "config": {"workload": "N=1024 cube scene fit to a second scene's silhouettes+depths, 16 views "
                                   "240x320, BCE + 0.1 L1, Adam lr 1e-2"}}


def workload_lone(fea, sc, dev, steps):
    """Latency of a lone request: one view per call, synchronised after each, N=4096 Gaussians
    in the unit cube seen from radius 3, forward and forward+backward, at 128x128, 240x320 and
    480x640 (f = W * 600/640), with 16-pixel tiles and with the server's tile rule (tile_for)."""
    import numpy as np
    import torch
    from fuzzy_ellipses_amd.serve import tile_for
    rng = np.random.default_rng(6)
    means, prec, w = sc.random_gaussians(rng, 4096)
    st = fea.SceneTensors.from_scene(sc.Scene(means, prec, w, []), dev)
    gen = torch.Generator(device="cpu").manual_seed(6)
    res = {}
    for H, W in ((128, 128), (240, 320), (480, 640)):
        cam = sc.make_camera(3.0 * sc.random_dir(rng), H, W, W * 600.0 / 640.0)
        gd = torch.randn(H, W, generator=gen).to(dev)
        ga = torch.randn(H, W, generator=gen).to(dev)
        for bwd in (False, True):
            for tile in sorted({16, tile_for(4096, H, W, bwd)}):
                r = fea.Renderer(tile=tile, device=dev)

                def call():
                    r.forward([st], [(0, cam, 21.4, 2.66)], save_for_backward=bwd)
                    if bwd:
                        r.backward([gd], [ga])
                for _ in range(3):
                    call()
                torch.cuda.synchronize()
                ts = []
                for _ in range(steps):
                    t0 = time.perf_counter()
                    call()
                    torch.cuda.synchronize()
                    ts.append(time.perf_counter() - t0)
                res[f"{H}x{W}_{'fwd_bwd' if bwd else 'fwd'}_tile{tile}_ms"] = round(1e3 * float(np.median(ts)), 4)
    return {"metric": "lone-request latency, median ms per synchronised call (N=4096)",
            "value": res["128x128_fwd_bwd_tile16_ms"], "unit": "ms", "n_gpus": 1, "steps": steps,
            "higher_is_better": False, "dtype": "f32", "data": "synthetic", "latency_ms": res,
            "config": {"workload": "one view per call of N=4096 Gaussians in the unit cube from radius 3; "
                                   "value = 128x128 forward+backward, 16-pixel tiles"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--tile", type=int, default=16)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--mode", default="culled")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", default="shell", choices=["shell", "batched", "serving", "silhouette", "gmm", "lone"])
    ap.add_argument("--rates", default="50,200,800")
    ap.add_argument("--requests", default="300,600,1200")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        src = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.add(src, 0.0, out=dst)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": src.numel() * 4, "bytes_written": dst.numel() * 4}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (views and scenes "
              f"shard across GPUs with no exchange, one process per GPU)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    fea = _package()
    from fuzzy_ellipses_amd import _lib
    import fuzzy_ellipses_amd.scenes as sc

    dev = torch.device("cuda:0")
    if args.workload == "batched":
        print(json.dumps(workload_batched(fea, sc, dev, args.steps)), flush=True)
        return
    if args.workload == "serving":
        rates = [float(x) for x in args.rates.split(",")]
        nreq = [int(x) for x in args.requests.split(",")]
        print(json.dumps(workload_serving(fea, sc, dev, rates, nreq)), flush=True)
        return
    if args.workload == "gmm":
        print(json.dumps(workload_gmm(fea, sc, dev, max(args.steps, 10))), flush=True)
        return
    if args.workload == "silhouette":
        print(json.dumps(workload_silhouette(fea, sc, dev, max(args.steps, 200))), flush=True)
        return
    if args.workload == "lone":
        print(json.dumps(workload_lone(fea, sc, dev, max(args.steps, 20))), flush=True)
        return
    scene = sc.shell_scene()
    cams = scene.cameras
    views = [(0, c, scene.beta2, scene.beta3) for c in cams]
    st = fea.SceneTensors.from_scene(scene, dev)
    r = fea.Renderer(tile=args.tile, batch=args.batch, mode=args.mode, device=dev)
    g = torch.Generator(device="cpu").manual_seed(0)
    gD = [torch.randn(c.H, c.W, generator=g).to(dev) for c in cams]
    gA = [torch.randn(c.H, c.W, generator=g).to(dev) for c in cams]

    def step():
        r.forward([st], views, save_for_backward=True)
        return r.backward(gD, gA)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    stats = r.stats()

    # per-kernel breakdown, every launch between two events, in a loop of its own (the events
    # add about 3% to a step, so the timed region below times only the roofline kernel)
    candidates = ("blend1", "blend2", "blend_dense", "bwd_blend")
    _lib.profile_read(reset=True)
    _lib.profile_enable(True)
    for _ in range(min(args.steps, 10)):
        step()
    torch.cuda.synchronize()
    _lib.profile_enable(False)
    kt_all = _lib.profile_read(reset=True)
    dom = max(candidates, key=lambda k: kt_all[k][0])   # the dominant kernel

    # timed region: forward + backward, with the dominant kernel's launches event-timed
    _lib.profile_read(reset=True)
    _lib.profile_enable(True, kernels=[dom])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    _lib.profile_enable(False)
    kt = _lib.profile_read(reset=True)

    # forward only: the median of three timed loops (a loop of 20 calls lasts ~25 ms, so one
    # host stall can double a single loop's time)
    for _ in range(2):
        r.forward([st], views)
    loops = []
    for _ in range(3):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            r.forward([st], views)
        torch.cuda.synchronize()
        loops.append(time.perf_counter() - t1)
    el_f = sorted(loops)[1]

    n_views = len(cams)
    P = cams[0].H * cams[0].W
    evals = stats["evals"]                 # evaluated pairs of one 16-view forward
    dense = scene.n * P * n_views          # dense-equivalent pairs
    fb_fps = n_views * args.steps / el
    f_fps = n_views * args.steps / el_f

    # roofline of the dominant kernel, its launches timed over the timed region
    pairs = {"blend1": stats["evals_phase1"], "blend2": evals - stats["evals_phase1"], "blend_dense": evals,
             "bwd_blend": evals}
    n_rec = scene.n * n_views
    K = stats["pairs_phase1"] + stats["pairs_phase2"]
    # algorithmic HBM bytes: per pixel 8 B out + 16 B saved aggregates (forward) or 16 B
    # aggregates + 8 B cotangents read (backward); 4 B per listed pair; per (Gaussian, view)
    # the backward reads the 64-B record and writes its 11 sums (44 B)
    alg_bytes = {"blend1": 24 * P * n_views + 4 * stats["pairs_phase1"],
                 "blend2": 24 * P * n_views + 4 * stats["pairs_phase2"],
                 "blend_dense": 24 * P * n_views,
                 "bwd_blend": 24 * P * n_views + 4 * K + 108 * n_rec}
    kavg = {k: (ms / n if n else 0.0) for k, (ms, n) in kt_all.items()}
    avg_s = (kt[dom][0] / kt[dom][1] if kt[dom][1] else 0.0) / 1e3
    roof = roofline(dom, avg_s, pairs[dom], FLOPS_BWD if dom == "bwd_blend" else FLOPS_FWD, alg_bytes[dom], args)
    fwd_k = "blend1" if args.mode == "culled" else "blend_dense"
    roof_fwd = roofline(fwd_k, kavg[fwd_k] / 1e3, pairs[fwd_k], FLOPS_FWD, alg_bytes[fwd_k], args)
    roof_fwd["timing"] = "every launch event-timed in the breakdown loop before the timed region"

    out = {
        "metric": METRIC, "value": round(fb_fps, 3), "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * el / args.steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "shell scene N=4096, 16 views 480x640 f=600, batched fwd+bwd per step",
                   "n_gaussians": scene.n, "views": n_views, "height": cams[0].H, "width": cams[0].W,
                   "tile": args.tile, "lds_batch": args.batch, "mode": args.mode},
        "fwd_frames_per_s": round(f_fps, 3),
        "fwd_evals_per_s": evals * args.steps / el_f,
        "fwd_dense_equiv_evals_per_s": dense * args.steps / el_f,
        "fwd_bwd_evals_per_s": 2 * evals * args.steps / el,
        "evals_per_forward": evals, "dense_pairs_per_forward": dense,
        "pairs_listed": stats["pairs_phase1"] + stats["pairs_phase2"], "floor_tiles": stats["floor_tiles"],
        "kernel_avg_ms": {k: round(v, 5) for k, v in kavg.items() if kt_all[k][1]},
        "roofline": roof,
        "roofline_forward_blend": roof_fwd,
    }
    if not args.no_batched:
        out["batched"] = batched_summary(fea, sc, dev, steps=max(args.steps // 2, 5))
    if not args.no_serving:
        rates = [float(x) for x in args.rates.split(",")]
        nreq = [int(x) for x in args.requests.split(",")]
        out["served"] = serving_summary(fea, sc, dev, rates, nreq)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(scene, cams[0])
        out["cpu_forward_f32"] = cpu_forward_f32(scene, cams[0])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
