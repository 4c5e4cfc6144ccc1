"""Throughput of DiffDVR's differentiable ray marcher on one MI355X.

Metric: forward+adjoint ray-samples/s of the batched, checkpointed path -- ONE
persistent launch renders every view, forms the image-loss gradient in
registers and runs the checkpointed reverse march -- against the plain
per-view form (per view: reference-structured forward launch, torch loss
gradient, reference-structured inversion adjoint launch).  A ray-sample is one
blended march iteration; N_samples = sum of the termination indices of the
forward pass, and the adjoint processes the same N.

Workloads (synthetic, seeded, identical on every host; diffdvr_amd/scenes.py):
  c5 (default: the largest single-GPU configuration): 16 volumes {64..256}^3 with
     4-32 cameras each at 128..512, linear TF per volume, L2, density + TF
     gradient, HIP Adam, 100 steps captured as one HIP graph; beside it the plain
     per-view form (reference-structured forward + inversion adjoint per view,
     then the same Adam step), the FP32 CPU port on sampled views, and the FP64
     oracle's error on those views
  c1: forward only, the CPU-runnable configuration: 64^3 blobs, 5-point linear TF,
     one 128^2 view at lat 30 / lon 45, step 0.5 voxel; the FP32 CPU port on the
     whole image beside it (1 warm-up, median of 5)
  c2: 128^3 blobs, 64x4 texture TF,
     8 views 256^2 at latitude 20, step 0.25 voxel, L1 vs a second TF, TF gradient
  c3: 256^3 blobs + noise, identity TF x10, 64 Fibonacci views 256^2, step 0.5
     voxel, L2 vs a second volume, gradient w.r.t. all voxels, K autotuned
  c4: 128^3 blobs, 4-Gaussian TF, 512 cameras of sizes {64..512}^2, L2 vs images
     from perturbed cameras, gradient w.r.t. latitude / longitude / radius
  fwd: forward mode on the C2 scene, D = 7 (per-pixel step size, ray start, ray
     direction), contracted with the L1 loss gradient
  preshade: TF pre-shading of a 256^3 volume (4-Gaussian TF) and its density
     adjoint, with best fit, log-MSE and importance sampling timed beside it

Prints ONE JSON line.  Options:
  --workload c1|c2|c3|c4|c5|fwd|preshade  --steps K --warmup W  --gpus N (only 1)
  --no-cpu-baseline  --profile (short run for rocprofv3: no plain leg, no baseline)
  --path persistent|fused (c2: persistent plan, or the single-batch fused kernel)
  --checkpoint-interval K  --wave-width W  --grid G  (all 0 = autotuned; c4 uses defaults)
  --pmc-probe (3 batched steps and one 256 MiB device copy, for counter calibration)
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12            # MI355X HBM3E, bytes/s
# algorithmic bytes per ray-sample of the batched kernel: 32 B of trilinear corner
# gathers in the forward march, 32 B again for the adjoint, plus 8 atomic
# read-modify-writes of 8 B for a volume gradient (TF and camera gradients are
# privatised in LDS / registers: 0 B per sample)
BYTES_PER_SAMPLE = {"c1": 32, "c2": 64, "c3": 128, "c4": 64, "c5": 128, "fwd": 32}
METRIC = "fwd+adjoint ray-samples/s on 1 MI355X, batched vs plain, at 1e-4 rel grad err"


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--workload", default="c5", choices=["c1", "c2", "c3", "c4", "c5", "fwd", "preshade"])
    p.add_argument("--steps", type=int, default=0,
                   help="timed steps (0 = per-workload default; c5: rounded up to whole 100-step graphs)")
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--profile", action="store_true")
    p.add_argument("--path", default="persistent", choices=["persistent", "fused"])
    p.add_argument("--checkpoint-interval", type=int, default=0)
    p.add_argument("--wave-width", type=int, default=0)
    p.add_argument("--grid", type=int, default=0, help="persistent workgroups (0 = resident capacity)")
    p.add_argument("--graph-steps", type=int, default=100)
    p.add_argument("--pmc-probe", action="store_true")
    return p.parse_args()


def traffic_from_profiles(name):
    """HBM bytes per launch of the batched kernel from a committed PMC summary, if present."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch="):
            try:
                return float(line.split("=", 1)[1])
            except ValueError:
                return None
    return None


def timed(fn, steps, stream):
    import torch
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def result_line(args, wl_name, value, ms, n_samples, config, kernel_ms, extra=None, kernel=None, metric=METRIC):
    achieved = n_samples * BYTES_PER_SAMPLE[wl_name] / (kernel_ms * 1e-3)
    if kernel is None:
        kernel = ("dvr::k_batched" if (wl_name != "c2" or args.path == "persistent")
                  else "dvr::k_adjoint_tiled (fused)")
    r = {
        "metric": metric, "value": value, "unit": "ray-samples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic", "config": config,
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK,
                     "traffic": traffic_from_profiles(f"pmc_traffic_batched_{wl_name.upper()}.txt"),
                     "kernel": kernel, "kernel_ms": kernel_ms, "bytes_per_sample": BYTES_PER_SAMPLE[wl_name]},
    }
    if extra:
        r.update(extra)
    return r


# ---------------------------------------------------------------------------
# plan construction shared by the workloads
# ---------------------------------------------------------------------------
def reference_frame_views(bt, cams_vp, sizes, vol_idx, tf_idx, fov, step, targets=None, grads=None):
    views = []
    for i, (w, h) in enumerate(sizes):
        views.append(bt.BatchView(volume=vol_idx, tf=tf_idx, width=w, height=h, camera=cams_vp[i], fov_y=fov,
                                  step=step, target=None if targets is None else targets[i],
                                  camera_grad=None if grads is None else grads[i]))
    return views


def render_targets(bt, vols, tfs, views, tf_mode, filter_mode=1):
    """Forward-only plan: images of (possibly different) volumes / cameras, outside any timing."""
    import torch
    plan = bt.BatchPlan([bt.BatchVolume(v) for v in vols], [bt.BatchTf(t) for t in tfs], views,
                        volume_filter_mode=filter_mode, tf_mode=tf_mode, forward_only=True)
    plan.run()
    torch.cuda.synchronize()
    return [v.color for v in views]


# ---------------------------------------------------------------------------
# C2 (the metric's configuration)
# ---------------------------------------------------------------------------
def make_inputs(pr, wl, tf, viewports=None):
    inp = pr.RendererInputs()
    inp.screen_size = pr.int2(wl.width, wl.height)
    inp.volume = wl.volume
    inp.volume_filter_mode = pr.VolumeFilterMode.Trilinear
    inp.box_min = pr.real3(-0.5, -0.5, -0.5)
    inp.box_size = pr.real3(1.0, 1.0, 1.0)
    inp.camera_mode = pr.CameraMode.ReferenceFrame
    inp.camera = pr.CameraReferenceFrame(wl.viewports if viewports is None else viewports, wl.fov_y)
    inp.step_size = wl.step
    inp.tf_mode = pr.TFMode.Texture
    inp.tf = tf
    inp.blend_mode = pr.BlendMode.BeerLambert
    return inp


def cpu_baseline_c2(wl, target, color, term, adj_tf, scale):
    """The FP32 OpenMP restatement of renderer_cpu.cpp on the whole workload
    (forward + L1 gradient + inversion adjoint with TF gradient), and the FP64
    error of the GPU's batched result (same sample counts)."""
    import numpy as np
    from oracle import oracle

    threads = min(16, os.cpu_count() or 1)
    oracle.set_threads(threads)
    s = dict(W=wl.width, H=wl.height, filter=1, camera=2, tf_mode=1, blend=0, early_out=1 - 1e-5, fov_y=wl.fov_y,
             volume=wl.volume.double().cpu().numpy(), box_min=np.full((1, 3), -0.5), box_size=np.ones((1, 3)),
             step=np.full((1, 1, 1), wl.step), tf=wl.tf.double().cpu().numpy(),
             cam_matrix=wl.viewports.double().cpu().numpy())
    tg = target.double().cpu().numpy()
    t0 = time.perf_counter()
    c32, t32 = oracle.forward(s, dtype=np.float32)
    adj = (np.sign(c32 - tg) * scale).astype(np.float32)
    oracle.adjoint(s, c32, t32, adj, grads=("tf",), dtype=np.float32)
    dt = time.perf_counter() - t0
    samples = float(t32.sum())
    c_gpu = color.double().cpu().numpy()
    t_gpu = term.cpu().numpy()
    c64, _ = oracle.forward(s, frozen_term=t_gpu)
    ferr = float(np.abs(c_gpu - c64).max() / np.abs(c64).max())
    g64 = oracle.adjoint(s, c64, t_gpu, np.sign(c_gpu - tg) * scale, grads=("tf",), method=oracle.EXACT_REPLAY)["tf"]
    g = adj_tf.double().cpu().numpy()
    gerr = float(np.linalg.norm(g - g64) / np.linalg.norm(g64))
    return {
        "value": samples / dt, "unit": "ray-samples/s", "cores": threads, "kind": "port",
        "sample": f"the whole workload once ({wl.B} views {wl.width}x{wl.height}, {samples:.3g} samples, {dt:.2f} s): "
                  f"FP32 forward + L1 gradient + inversion adjoint with TF gradient, the OpenMP restatement of "
                  f"renderer_cpu.cpp (schedule(guided) over pixels, omp atomic scatters)",
    }, ferr, gerr


# ---------------------------------------------------------------------------
# C1: forward only (the CPU-runnable configuration)
# ---------------------------------------------------------------------------
def bench_c1(args):
    import numpy as np
    import torch
    from diffdvr_amd import pyrenderer as pr
    from diffdvr_amd import scenes
    from oracle import oracle

    dev = torch.device("cuda:0")
    wl = scenes.workload_c1().to(dev)
    H, W = wl.height, wl.width
    steps = args.steps or 200
    args.steps = steps
    stream = torch.cuda.current_stream()
    inp = pr.RendererInputs()
    inp.screen_size = pr.int2(W, H)
    inp.volume = wl.volume
    inp.volume_filter_mode = pr.VolumeFilterMode.Trilinear
    inp.box_min = pr.real3(-0.5, -0.5, -0.5)
    inp.box_size = pr.real3(1.0, 1.0, 1.0)
    inp.camera_mode = pr.CameraMode.ReferenceFrame
    inp.camera = pr.CameraReferenceFrame(wl.viewports, wl.fov_y)
    inp.step_size = wl.step
    inp.tf_mode = pr.TFMode.Linear
    inp.tf = wl.tf
    inp.blend_mode = pr.BlendMode.BeerLambert
    color = torch.empty(1, H, W, 4, device=dev)
    term = torch.empty(1, H, W, dtype=torch.int32, device=dev)
    outs = pr.RendererOutputs(color, term)

    def tiled():
        pr.Renderer.render_forward(inp, outs, path="tiled")

    def plain():
        pr.Renderer.render_forward(inp, outs, path="plain")

    for _ in range(args.warmup):
        tiled()
        plain()
    torch.cuda.synchronize()
    ms_plain = timed(plain, steps, stream)
    ms = timed(tiled, steps, stream)
    n = float(term.sum().item())
    config = {"workload": "C1: 64^3 blob volume (8 blobs), 5-point linear TF, one 128x128 view (lat 30, lon 45, "
                          "radius 2.5), step 0.5 voxel, forward only", "image": [H, W], "samples_per_step": n,
              "path": "tiled"}
    result = result_line(args, "c1", n / (ms * 1e-3), ms, n, config, ms, kernel="dvr::k_forward_tiled",
                         metric="forward ray-samples/s on 1 MI355X")
    result["plain"] = {"value": n / (ms_plain * 1e-3), "ms_per_step": ms_plain, "speedup_tiled_vs_plain": ms_plain / ms}
    s = dict(W=W, H=H, filter=1, camera=2, tf_mode=2, blend=0, early_out=1 - 1e-5, fov_y=wl.fov_y,
             volume=wl.volume.double().cpu().numpy(), box_min=np.full((1, 3), -0.5), box_size=np.ones((1, 3)),
             step=np.full((1, 1, 1), wl.step), tf=wl.tf.double().cpu().numpy(),
             cam_matrix=wl.viewports.double().cpu().numpy())
    c64, _ = oracle.forward(s, frozen_term=term.cpu().numpy())
    result["accuracy"] = {"forward_rel_err_vs_fp64": float(np.abs(color.double().cpu().numpy() - c64).max() /
                                                           np.abs(c64).max())}
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        oracle.set_threads(threads)
        oracle.forward(s, dtype=np.float32)
        times = []
        for _ in range(5):
            t0 = time.perf_counter()
            _, t32 = oracle.forward(s, dtype=np.float32)
            times.append(time.perf_counter() - t0)
        med = float(np.median(times))
        result["cpu_baseline"] = {
            "value": float(t32.sum()) / med, "unit": "ray-samples/s", "cores": threads, "kind": "port",
            "sample": f"the whole workload ({W}x{H}, {float(t32.sum()):.3g} samples): FP32 forward, the OpenMP "
                      f"restatement of renderer_cpu.cpp; 1 warm-up, median of 5 ({med * 1e3:.2f} ms)",
            **host_cpu()}
    return result


def bench_c2(args):
    import numpy as np
    import torch
    from diffdvr_amd import pyrenderer as pr
    from diffdvr_amd import scenes
    from diffdvr_amd import batch as bt

    dev = torch.device("cuda:0")
    wl = scenes.workload_c2(device="cpu").to(dev)
    B, H, W = wl.B, wl.height, wl.width
    stream = torch.cuda.current_stream()
    steps = args.steps or 20
    args.steps = steps

    target = torch.empty(B, H, W, 4, device=dev)
    tterm = torch.empty(B, H, W, dtype=torch.int32, device=dev)
    pr.Renderer.render_forward(make_inputs(pr, wl, wl.target_tf), pr.RendererOutputs(target, tterm))

    adj_tf = torch.zeros_like(wl.tf)
    if args.path == "persistent":
        vol = wl.volume[0].contiguous()
        tf = wl.tf[0]
        color = torch.empty(B, H, W, 4, device=dev)
        term = torch.empty(B, H, W, dtype=torch.int32, device=dev)
        views = [bt.BatchView(volume=0, tf=0, width=W, height=H, camera=wl.viewports[b], fov_y=wl.fov_y,
                              step=wl.step, color=color[b], termination=term[b], target=target[b])
                 for b in range(B)]
        plan = bt.BatchPlan([bt.BatchVolume(vol)], [bt.BatchTf(tf, adj_tf[0])], views, volume_filter_mode=1,
                            tf_mode=1, blend_mode=0, loss="l1",
                            checkpoint_interval=args.checkpoint_interval or 16, wave_width=args.wave_width or 8,
                            grid=args.grid)

        if not (args.checkpoint_interval or args.wave_width or args.grid):
            plan.autotune(iterations=3)

        def batched_step():
            plan.run(zero_gradients=True)
        options = plan.options
    else:
        inp = make_inputs(pr, wl, wl.tf)
        color = torch.empty(B, H, W, 4, device=dev)
        term = torch.empty(B, H, W, dtype=torch.int32, device=dev)
        adj = pr.AdjointOutputs()
        adj.has_tf_derivatives = True
        adj.adj_tf = adj_tf
        outs = pr.RendererOutputs(color, term)

        def batched_step():
            adj_tf.zero_()
            pr.Renderer.render_fused(inp, outs, "l1", target, adj, checkpoint_interval=args.checkpoint_interval)
        options = {"checkpoint_interval": args.checkpoint_interval or 16}

    if args.pmc_probe:
        for _ in range(3):
            batched_step()
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=dev)   # 256 MiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"pmc_probe": True, "copy_bytes": src.numel() * 4,
                          "samples_per_step": float(term.sum().item())}))
        return None

    for _ in range(args.warmup):
        batched_step()
    torch.cuda.synchronize()
    n_samples = float(term.sum().item())
    ms_batched = timed(batched_step, steps, stream)
    value = n_samples / (ms_batched * 1e-3)
    if args.path == "persistent":   # the persistent kernel alone (HIP events around it on its stream)
        kernel_ms = float(np.median([plan.kernel_ms() for _ in range(max(3, steps // 2))]))
    else:   # one call = the gradient clear and the fused kernel
        kernel_ms = float(np.mean([timed(batched_step, 1, stream) for _ in range(max(3, steps // 2))]))

    config = {"workload": "C2: 128^3 blob volume, 64x4 texture TF gradient, 8 views 256x256, step 0.25 voxel, "
                          "L1 loss, one persistent fwd+loss+checkpointed-adjoint launch"
                          if args.path == "persistent" else
                          "C2: 128^3 blob volume, 64x4 texture TF gradient, 8 views 256x256, step 0.25 voxel, "
                          "L1 loss, single-batch fused fwd+loss+checkpointed-adjoint launch",
              "views": B, "image": [H, W], "volume": list(wl.volume.shape[1:]), "samples_per_step": n_samples,
              "path": args.path, **options}
    result = result_line(args, "c2", value, ms_batched, n_samples, config, kernel_ms)
    if args.path == "fused":
        result["roofline"]["traffic"] = traffic_from_profiles("pmc_traffic_fused_C2.txt")

    if not args.profile:
        views_p = []
        for b in range(B):
            ib = make_inputs(pr, wl, wl.tf, wl.viewports[b:b + 1].contiguous())
            cb = torch.empty(1, H, W, 4, device=dev)
            tb = torch.empty(1, H, W, dtype=torch.int32, device=dev)
            views_p.append((ib, cb, tb, target[b:b + 1].contiguous()))
        padj_tf = torch.zeros_like(wl.tf)
        padj = pr.AdjointOutputs()
        padj.has_tf_derivatives = True
        padj.tf_delayed_accumulation = True   # diffdvr.Renderer default (pytests/diffdvr/renderer.py:270)
        padj.adj_tf = padj_tf
        n_all = float(B * H * W * 4)

        def plain_step():
            padj_tf.zero_()
            for ib, cb, tb, tgb in views_p:
                o = pr.RendererOutputs(cb, tb)
                pr.Renderer.render_forward(ib, o, path="plain")
                g = torch.sign(cb - tgb) / n_all
                pr.Renderer.render_adjoint(ib, o, g, padj, method="inversion", path="plain")

        for _ in range(max(1, args.warmup)):
            plain_step()
        ms_plain = timed(plain_step, max(3, steps // 4), stream)
        plain_value = n_samples / (ms_plain * 1e-3)
        batched_step()
        torch.cuda.synchronize()
        plain_err = float(torch.linalg.norm(padj_tf - adj_tf) / torch.linalg.norm(adj_tf))
        result["plain"] = {"value": plain_value, "ms_per_step": ms_plain, "speedup_batched_vs_plain": value / plain_value,
                           "tf_grad_rel_diff_vs_batched": plain_err}
        if not args.no_cpu_baseline:
            base, ferr, gerr = cpu_baseline_c2(wl, target, color, term, adj_tf, 1.0 / n_all)
            result["cpu_baseline"] = base
            result["accuracy"] = {"forward_rel_err_vs_fp64": ferr, "tf_grad_rel_err_vs_fp64": gerr}
    return result


# ---------------------------------------------------------------------------
# C3: volume gradient
# ---------------------------------------------------------------------------
def bench_c3(args):
    import numpy as np
    import torch
    from diffdvr_amd import pyrenderer as pr
    from diffdvr_amd import scenes
    from diffdvr_amd import batch as bt

    dev = torch.device("cuda:0")
    wl = scenes.workload_c3(device="cpu").to(dev)
    B, H, W = wl.B, wl.height, wl.width
    steps = args.steps or 5
    args.steps = steps
    stream = torch.cuda.current_stream()
    vol = wl.volume[0].contiguous()
    tf = wl.tf[0].contiguous()
    vps = [wl.viewports[b] for b in range(B)]
    tviews = reference_frame_views(bt, vps, [(W, H)] * B, 0, 0, wl.fov_y, wl.step)
    targets = render_targets(bt, [wl.target_volume[0].contiguous()], [tf], tviews, 0)
    adj_vol = torch.zeros_like(vol)
    views = reference_frame_views(bt, vps, [(W, H)] * B, 0, 0, wl.fov_y, wl.step, targets=targets)
    plan = bt.BatchPlan([bt.BatchVolume(vol, grad=adj_vol)], [bt.BatchTf(tf)], views, volume_filter_mode=1,
                        tf_mode=0, blend_mode=0, loss="l2",
                        checkpoint_interval=args.checkpoint_interval or 16, wave_width=args.wave_width or 8,
                        grid=args.grid)
    tuned = None
    if not (args.checkpoint_interval or args.wave_width or args.grid):
        tuned = plan.autotune(iterations=1)

    def step():
        plan.run(zero_gradients=True)

    if args.pmc_probe:   # 3 batched runs and one 256 MiB device copy (counter calibration)
        for _ in range(3):
            step()
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"pmc_probe": True, "copy_bytes": src.numel() * 4,
                          "samples_per_step": float(sum(v.termination.sum().item() for v in views))}))
        return None

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    n_samples = float(sum(v.termination.sum().item() for v in views))
    ms = timed(step, steps, stream)
    value = n_samples / (ms * 1e-3)
    config = {"workload": "C3: 256^3 blob volume + noise 0.02, identity TF x10, 64 Fibonacci views 256x256, "
                          "step 0.5 voxel, L2 vs a second volume, gradient w.r.t. all voxels, K autotuned",
              "views": B, "image": [H, W], "samples_per_step": n_samples, **plan.options,
              "chunk_slots": plan.chunk_slots}
    # the persistent kernel alone (HIP events around it on the launch stream)
    k_ms = float(np.median([plan.kernel_ms() for _ in range(5)]))
    result = result_line(args, "c3", value, ms, n_samples, config, k_ms,
                         {"autotuned": tuned is not None})
    if not args.profile:
        inp = pr.RendererInputs()
        inp.screen_size = pr.int2(W, H)
        inp.volume = wl.volume
        inp.volume_filter_mode = pr.VolumeFilterMode.Trilinear
        inp.box_min = pr.real3(-0.5, -0.5, -0.5)
        inp.box_size = pr.real3(1, 1, 1)
        inp.camera_mode = pr.CameraMode.ReferenceFrame
        inp.step_size = wl.step
        inp.tf_mode = pr.TFMode.Identity
        inp.tf = wl.tf
        inp.blend_mode = pr.BlendMode.BeerLambert
        padj_vol = torch.zeros_like(wl.volume)
        padj = pr.AdjointOutputs()
        padj.has_volume_derivatives = True
        padj.adj_volume = padj_vol
        n_all = float(B * H * W * 4)
        per = []
        for b in range(B):
            cb = torch.empty(1, H, W, 4, device=dev)
            tb = torch.empty(1, H, W, dtype=torch.int32, device=dev)
            per.append((pr.CameraReferenceFrame(wl.viewports[b:b + 1].contiguous(), wl.fov_y), cb, tb,
                        targets[b].unsqueeze(0)))

        def plain_step():
            padj_vol.zero_()
            for cam, cb, tb, tgb in per:
                inp.camera = cam
                o = pr.RendererOutputs(cb, tb)
                pr.Renderer.render_forward(inp, o, path="plain")
                pr.Renderer.render_adjoint(inp, o, 2.0 * (cb - tgb) / n_all, padj, method="inversion", path="plain")

        plain_step()
        ms_plain = timed(plain_step, max(2, steps // 2), stream)
        step()
        torch.cuda.synchronize()
        diff = float(torch.linalg.norm(padj_vol[0] - adj_vol) / torch.linalg.norm(adj_vol))
        result["plain"] = {"value": n_samples / (ms_plain * 1e-3), "ms_per_step": ms_plain,
                           "speedup_batched_vs_plain": ms_plain / ms, "vol_grad_rel_diff_vs_batched": diff}
    return result


# ---------------------------------------------------------------------------
# C4: camera gradients of 512 cameras of uneven size
# ---------------------------------------------------------------------------
def bench_c4(args):
    import numpy as np
    import torch
    from diffdvr_amd import pyrenderer as pr
    from diffdvr_amd import scenes
    from diffdvr_amd import batch as bt

    dev = torch.device("cuda:0")
    wl = scenes.workload_c4()
    cams, tcams = wl.cameras[0], wl.target_cameras[0]
    N = len(cams.sizes)
    steps = args.steps or 5
    args.steps = steps
    stream = torch.cuda.current_stream()
    vol = wl.volumes[0].to(dev).contiguous()
    tf = wl.tfs[0].to(dev).contiguous()
    step_w = wl.steps[0]
    tvp = tcams.viewports(dev)
    tviews = reference_frame_views(bt, [tvp[i] for i in range(N)], cams.sizes, 0, 0, wl.fov_y, step_w)
    targets = render_targets(bt, [vol], [tf], tviews, 3)
    del tviews
    lat = cams.lat.to(dev).float().requires_grad_(True)
    lon = cams.lon.to(dev).float().requires_grad_(True)
    rad = cams.radius.to(dev).float().requires_grad_(True)
    vp_buf = torch.zeros(N, 3, 3, device=dev)
    g_vp = torch.zeros(N, 3, 3, device=dev)
    views = reference_frame_views(bt, [vp_buf[i] for i in range(N)], cams.sizes, 0, 0, wl.fov_y, step_w,
                                  targets=targets, grads=[g_vp[i] for i in range(N)])
    loss = torch.zeros(1, device=dev)
    plan = bt.BatchPlan([bt.BatchVolume(vol)], [bt.BatchTf(tf)], views, volume_filter_mode=1, tf_mode=3,
                        blend_mode=0, loss="l2", loss_out=loss,
                        checkpoint_interval=args.checkpoint_interval or 16, wave_width=args.wave_width or 8,
                        grid=args.grid)

    def step():
        for p in (lat, lon, rad):
            p.grad = None
        vp = pr.Camera.viewport_from_sphere(torch.zeros(N, 3, device=dev), lon, lat, rad, pr.Orientation.Yp)
        vp_buf.copy_(vp.detach())
        plan.run(zero_gradients=True)
        vp.backward(g_vp)

    tuned = None
    if not (args.checkpoint_interval or args.wave_width or args.grid):
        with torch.no_grad():  # the plan reads the cameras from vp_buf
            vp_buf.copy_(pr.Camera.viewport_from_sphere(torch.zeros(N, 3, device=dev), lon, lat, rad,
                                                        pr.Orientation.Yp))
        tuned = plan.autotune(iterations=1)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    n_samples = float(sum(v.termination.sum().item() for v in views))
    ms = timed(step, steps, stream)
    k_ms = float(np.median([plan.kernel_ms() for _ in range(5)]))
    value = n_samples / (ms * 1e-3)
    config = {"workload": "C4: 128^3 blob volume, 4-component Gaussian TF, 512 sphere cameras of sizes "
                          "{64,128,256,512}^2, step 0.5 voxel, L2 vs perturbed-camera images, gradient w.r.t. "
                          "latitude/longitude/radius, one persistent launch",
              "cameras": N, "rays": int(sum(w * h for w, h in cams.sizes)), "samples_per_step": n_samples,
              **plan.options, "chunk_slots": plan.chunk_slots}
    result = result_line(args, "c4", value, ms, n_samples, config, k_ms, {"autotuned": tuned is not None})
    if not args.profile:
        inp = pr.RendererInputs()
        inp.volume = vol.unsqueeze(0)
        inp.volume_filter_mode = pr.VolumeFilterMode.Trilinear
        inp.box_min = pr.real3(-0.5, -0.5, -0.5)
        inp.box_size = pr.real3(1, 1, 1)
        inp.camera_mode = pr.CameraMode.RayStartDir
        inp.step_size = step_w
        inp.tf_mode = pr.TFMode.Gaussian
        inp.tf = tf.unsqueeze(0)
        inp.blend_mode = pr.BlendMode.BeerLambert
        n_all = float(sum(w * h for w, h in cams.sizes) * 4)
        bufs = []
        for i, (w, h) in enumerate(cams.sizes):
            rs = torch.empty(1, h, w, 3, device=dev)
            rd = torch.empty(1, h, w, 3, device=dev)
            ars = torch.zeros(1, h, w, 3, device=dev)
            ard = torch.zeros(1, h, w, 3, device=dev)
            cb = torch.empty(1, h, w, 4, device=dev)
            tb = torch.empty(1, h, w, dtype=torch.int32, device=dev)
            bufs.append((w, h, rs, rd, ars, ard, cb, tb, targets[i].unsqueeze(0)))
        pg_vp = torch.zeros(N, 3, 3, device=dev)

        def plain_step():
            for p in (lat, lon, rad):
                p.grad = None
            vp = pr.Camera.viewport_from_sphere(torch.zeros(N, 3, device=dev), lon, lat, rad, pr.Orientation.Yp)
            vpd = vp.detach()
            for i, (w, h, rs, rd, ars, ard, cb, tb, tgb) in enumerate(bufs):
                v1 = vpd[i:i + 1]
                pr._n.generate_rays(v1, wl.fov_y, w, h, rs, rd)
                inp.screen_size = pr.int2(w, h)
                inp.camera = pr.CameraPerPixelRays(rs, rd)
                o = pr.RendererOutputs(cb, tb)
                pr.Renderer.render_forward(inp, o, path="plain")
                a = pr.AdjointOutputs()
                a.has_camera_derivatives = True
                ars.zero_()
                ard.zero_()
                a.adj_camera_ray_start = ars
                a.adj_camera_ray_dir = ard
                pr.Renderer.render_adjoint(inp, o, 2.0 * (cb - tgb) / n_all, a, method="inversion", path="plain")
                pr._n.generate_rays_adjoint(v1, wl.fov_y, w, h, ars, ard, pg_vp[i:i + 1])
            vp.backward(pg_vp)

        plain_step()
        ms_plain = timed(plain_step, 2, stream)
        result["plain"] = {"value": n_samples / (ms_plain * 1e-3), "ms_per_step": ms_plain,
                           "speedup_batched_vs_plain": ms_plain / ms}
    return result


# ---------------------------------------------------------------------------
# C5: multi-volume optimisation loop, 100 steps as one HIP graph (the default
# workload: the largest single-GPU configuration)
# ---------------------------------------------------------------------------
def host_cpu():
    """CPU model, sockets and logical CPUs of this host (from /proc/cpuinfo)."""
    model, sockets, logical = "unknown", set(), 0
    try:
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k = k.strip()
            if k == "model name":
                model = v.strip()
                logical += 1
            elif k == "physical id":
                sockets.add(v.strip())
    except OSError:
        pass
    return {"cpu_model": model, "sockets": max(1, len(sockets)), "logical_cpus": logical or (os.cpu_count() or 0)}


def c5_oracle_scene(vol_np, tf_np, vp_np, W, H, step, fov):
    import numpy as np
    return dict(W=W, H=H, filter=1, camera=2, tf_mode=2, blend=0, early_out=1 - 1e-5, fov_y=fov,
                volume=vol_np[None], box_min=np.full((1, 3), -0.5), box_size=np.ones((1, 3)),
                step=np.full((1, 1, 1), step), tf=tf_np[None], cam_matrix=vp_np[None])


def cpu_baseline_c5(sample, n_all):
    """The FP32 OpenMP restatement of renderer_cpu.cpp (forward + L2 gradient + inversion
    adjoint with density and TF gradients) on the sampled views: one warm-up pass, then the
    median of five."""
    import numpy as np
    from oracle import oracle

    # one GPU's share of the host's cores: the box gives each GPU 16, whatever nproc shows
    threads = min(16, os.cpu_count() or 1)
    oracle.set_threads(threads)
    scenes32 = [(c5_oracle_scene(*(a.astype(np.float32) for a in (v, t, vp)), W, H, st, fov), tg)
                for (v, t, vp, W, H, st, fov, tg) in sample]

    def one_pass():
        n = 0.0
        for s, tg in scenes32:
            c, t = oracle.forward(s, dtype=np.float32)
            adj = (2.0 * (c - tg[None]) / n_all).astype(np.float32)
            oracle.adjoint(s, c, t, adj, grads=("volume", "tf"), dtype=np.float32)
            n += float(t.sum())
        return n

    one_pass()
    times, n = [], 0.0
    for _ in range(5):
        t0 = time.perf_counter()
        n = one_pass()
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    return {"value": n / med, "unit": "ray-samples/s", "cores": threads, "kind": "port",
            "sample": f"{len(sample)} of the workload's views (the first view of each volume in order until "
                      f"2.5e7 GPU samples and 3 views; {n:.3g} samples per pass): FP32 forward + L2 gradient + inversion "
                      f"adjoint with density and TF gradients, the OpenMP restatement of renderer_cpu.cpp "
                      f"(schedule(guided) over pixels, omp atomic scatters); 1 warm-up, median of 5 passes "
                      f"({med:.2f} s)",
            **host_cpu(), "threads_note": "16 threads = one GPU's share of the GPU host's cores"}


def bench_c5(args):
    import numpy as np
    import torch
    from diffdvr_amd import pyrenderer as pr
    from diffdvr_amd import scenes
    from diffdvr_amd import batch as bt

    dev = torch.device("cuda:0")
    wl = scenes.workload_c5(device=dev)
    G = args.graph_steps
    reps = max(1, -(-(args.steps or G) // G))
    args.steps = reps * G
    stream = torch.cuda.current_stream()
    vols = [v.contiguous() for v in wl.volumes]
    tgts = [v.contiguous() for v in wl.target_volumes]
    tfs = [t.to(dev).contiguous() for t in wl.tfs]
    vps, sizes, vidx, steps_w = [], [], [], []
    for k, cams in enumerate(wl.cameras):
        vp = cams.viewports(dev)
        for i in range(len(cams.sizes)):
            vps.append(vp[i].contiguous())
            sizes.append(cams.sizes[i])
            vidx.append(k)
            steps_w.append(wl.steps[k])

    def views_for(targets=None):
        return [bt.BatchView(volume=vidx[i], tf=vidx[i], width=sizes[i][0], height=sizes[i][1], camera=vps[i],
                             fov_y=wl.fov_y, step=steps_w[i], target=None if targets is None else targets[i])
                for i in range(len(vps))]

    targets = render_targets(bt, tgts, tfs, views_for(), 2)
    del tgts
    g_vols = [torch.zeros_like(v) for v in vols]
    g_tfs = [torch.zeros_like(t) for t in tfs]
    views = views_for(targets)
    loss = torch.zeros(1, device=dev)
    plan = bt.BatchPlan([bt.BatchVolume(v, grad=g) for v, g in zip(vols, g_vols)],
                        [bt.BatchTf(t, g) for t, g in zip(tfs, g_tfs)], views, volume_filter_mode=1, tf_mode=2,
                        blend_mode=0, loss="l2", loss_out=loss,
                        checkpoint_interval=args.checkpoint_interval or 16, wave_width=args.wave_width or 8,
                        grid=args.grid)
    tuned = None
    if not (args.checkpoint_interval or args.wave_width or args.grid):
        tuned = plan.autotune(iterations=1)
    n_all = float(sum(w * h for w, h in sizes) * 4)

    if args.pmc_probe:   # 3 batched runs and one 256 MiB device copy (counter calibration)
        for _ in range(3):
            plan.run()
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"pmc_probe": True, "copy_bytes": src.numel() * 4,
                          "samples_per_step": float(sum(v.termination.sum().item() for v in views))}))
        return None

    # accuracy sample: the first view of each volume until 2.5e7 GPU samples (also the CPU baseline's)
    plan.run()
    torch.cuda.synchronize()
    first = {}
    for i in range(len(views)):
        first.setdefault(vidx[i], i)
    sample_idx, acc_n = [], 0.0
    for k in sorted(first):
        if acc_n >= 2.5e7 and len(sample_idx) >= 3:
            break
        i = first[k]
        sample_idx.append(i)
        acc_n += float(views[i].termination.sum().item())
    accuracy, cpu_sample = None, None
    if not args.profile:
        accuracy, cpu_sample = c5_accuracy(bt, wl, vols, tfs, vps, sizes, vidx, steps_w, targets, sample_idx, n_all)

    # plain leg: per view the reference-structured forward, torch's L2 gradient and the inversion
    # adjoint (density + TF gradients accumulated into the same tensors), then the same Adam step
    adam = bt.Adam([(v, g, None, (0.0, 1.0)) for v, g in zip(vols, g_vols)] +
                   [(t, g, 4, (0.0, 1e3)) for t, g in zip(tfs, g_tfs)], lr=1e-2)
    plain = None
    if not args.profile:
        per = []
        for i in range(len(views)):
            w, h = sizes[i]
            inp = pr.RendererInputs()
            inp.screen_size = pr.int2(w, h)
            inp.volume = vols[vidx[i]].unsqueeze(0)
            inp.volume_filter_mode = pr.VolumeFilterMode.Trilinear
            inp.box_min = pr.real3(-0.5, -0.5, -0.5)
            inp.box_size = pr.real3(1.0, 1.0, 1.0)
            inp.camera_mode = pr.CameraMode.ReferenceFrame
            inp.camera = pr.CameraReferenceFrame(vps[i].unsqueeze(0), wl.fov_y)
            inp.step_size = steps_w[i]
            inp.tf_mode = pr.TFMode.Linear
            inp.tf = tfs[vidx[i]].unsqueeze(0)
            inp.blend_mode = pr.BlendMode.BeerLambert
            a = pr.AdjointOutputs()
            a.has_volume_derivatives = True
            a.adj_volume = g_vols[vidx[i]].unsqueeze(0)
            a.has_tf_derivatives = True
            a.tf_delayed_accumulation = True   # diffdvr.Renderer default (pytests/diffdvr/renderer.py:270)
            a.adj_tf = g_tfs[vidx[i]].unsqueeze(0)
            cb = torch.empty(1, h, w, 4, device=dev)
            tb = torch.empty(1, h, w, dtype=torch.int32, device=dev)
            per.append((inp, a, cb, tb, targets[i].unsqueeze(0)))

        def plain_grads():
            for g in g_vols + g_tfs:
                g.zero_()
            for inp, a, cb, tb, tg in per:
                o = pr.RendererOutputs(cb, tb)
                pr.Renderer.render_forward(inp, o, path="plain")
                pr.Renderer.render_adjoint(inp, o, 2.0 * (cb - tg) / n_all, a, method="inversion", path="plain")

        # gradients of both forms at the same parameters
        plan.run()
        torch.cuda.synchronize()
        bg = [g.clone() for g in g_vols[:1] + g_tfs]
        plain_grads()
        torch.cuda.synchronize()
        pg = [g.clone() for g in g_vols[:1] + g_tfs]
        d_vol = float(torch.linalg.norm(pg[0] - bg[0]) / torch.linalg.norm(bg[0]))
        d_tf = float(torch.linalg.norm(torch.cat([(a - b).flatten() for a, b in zip(pg[1:], bg[1:])])) /
                     torch.linalg.norm(torch.cat([b.flatten() for b in bg[1:]])))
        del bg, pg

        def plain_step():
            plain_grads()
            adam.step()

        plain_step()
        ms_plain = timed(plain_step, 2, stream)
        plain = {"ms_per_step": ms_plain, "density_grad_rel_diff_vs_batched": d_vol,
                 "tf_grad_rel_diff_vs_batched": d_tf}

    for _ in range(args.warmup):
        plan.run()
        adam.step()
    torch.cuda.synchronize()
    n0 = float(sum(v.termination.sum().item() for v in views))
    loss0 = float(loss.item())
    graph = bt.GraphedSteps(plan, adam, G)
    torch.cuda.synchronize()
    ms_graph = timed(graph.replay, reps, stream)
    n1 = float(sum(v.termination.sum().item() for v in views))   # the last step's sample count
    n_samples = 0.5 * (n0 + n1)
    ms = ms_graph / G
    value = n_samples / (ms * 1e-3)
    loss1 = float(loss.item())
    # the persistent kernel alone (HIP events around it on the launch stream), outside the graph
    k_ms = float(np.median([plan.kernel_ms() for _ in range(5)]))
    config = {"workload": "C5: 16 blob volumes {64..256}^3, 4-32 cameras each at 128..512, linear TF per volume, "
                          "step 0.5 voxel, L2 vs second volumes, density + TF gradient, HIP Adam, "
                          f"{G} steps captured as one HIP graph",
              "volumes": len(vols), "views": len(views), "rays": int(sum(w * h for w, h in sizes)),
              "samples_per_step": n_samples, "samples_first_last": [n0, n1], "graph_steps": G, "graph_replays": reps,
              "loss_before_graph": loss0, "loss_last_step": loss1, "adam_steps": adam.step_count, **plan.options,
              "chunk_slots": plan.chunk_slots}
    result = result_line(args, "c5", value, ms, n_samples, config, k_ms, {"autotuned": tuned is not None})
    if plain is not None:
        plain["value"] = n_samples / (plain["ms_per_step"] * 1e-3)
        plain["speedup_batched_vs_plain"] = plain["ms_per_step"] / ms
        result["plain"] = plain
    if accuracy is not None:
        result["accuracy"] = accuracy
        if not args.no_cpu_baseline:
            result["cpu_baseline"] = cpu_baseline_c5(cpu_sample, n_all)
    return result


def c5_accuracy(bt, wl, vols, tfs, vps, sizes, vidx, steps_w, targets, idx, n_all):
    """FP64 oracle vs the batched kernel on the sampled views: a plan of those views alone (same
    kernel instance) gives their gradients; colours at the GPU's sample counts, density (interior
    voxels) and TF gradients of the exact-replay adjoint.  The headline gradient errors leave out
    the rays with a sample within FP32 rounding of a control point (the oracle's kink census),
    whose loss gradient is zeroed on both sides; the all-ray errors and the FP32 CPU twin's
    all-ray density error are given beside them."""
    import numpy as np
    import torch
    from oracle import oracle

    oracle.set_threads(min(16, os.cpu_count() or 1))
    used = sorted({vidx[i] for i in idx})
    g_v = {k: torch.zeros_like(vols[k]) for k in used}
    g_t = {k: torch.zeros_like(tfs[k]) for k in used}
    tg_sub = [targets[i].clone() for i in idx]
    sub = [bt.BatchView(volume=used.index(vidx[i]), tf=used.index(vidx[i]), width=sizes[i][0], height=sizes[i][1],
                        camera=vps[i], fov_y=wl.fov_y, step=steps_w[i], target=tg_sub[j], loss_scale=1.0 / n_all)
           for j, i in enumerate(idx)]
    p = bt.BatchPlan([bt.BatchVolume(vols[k], grad=g_v[k]) for k in used], [bt.BatchTf(tfs[k], g_t[k]) for k in used],
                     sub, volume_filter_mode=1, tf_mode=2, blend_mode=0, loss="l2")
    p.run()
    torch.cuda.synchronize()
    got_all = {k: (g_v[k].double().cpu().numpy(), g_t[k].double().cpu().numpy()) for k in used}
    ferr, sample, kink_samples, samples = 0.0, [], 0, 0
    want = {k: {n: 0.0 for n in ("v", "t", "v_clean", "t_clean", "v_twin")} for k in used}
    for j, i in enumerate(idx):
        k = vidx[i]
        w, h = sizes[i]
        vol = vols[k].double().cpu().numpy()
        tf = tfs[k].double().cpu().numpy()
        vp = vps[i].double().cpu().numpy()
        tg = targets[i].double().cpu().numpy()
        s = c5_oracle_scene(vol, tf, vp, w, h, steps_w[i], wl.fov_y)
        t = sub[j].termination.cpu().numpy()[None]
        c64, _ = oracle.forward(s, frozen_term=t)
        c = sub[j].color.double().cpu().numpy()[None]
        ferr = max(ferr, float(np.abs(c - c64).max() / np.abs(c64).max()))
        adj = 2.0 * (c - tg[None]) / n_all
        g = oracle.adjoint(s, c64, t, adj, grads=("volume", "tf"), method=oracle.EXACT_REPLAY)
        want[k]["v"] = want[k]["v"] + g["volume"][0]
        want[k]["t"] = want[k]["t"] + g["tf"][0]
        c32, _ = oracle.forward(s, dtype=np.float32, frozen_term=t)
        want[k]["v_twin"] = want[k]["v_twin"] + oracle.adjoint(s, c32, t, adj, grads=("volume",), dtype=np.float32,
                                                                method=oracle.EXACT_REPLAY)["volume"][0]
        census = oracle.kink_census(s, t)[0]
        kink_samples += int(census.sum())
        samples += int(t.sum())
        kinked = census > 0
        kt = torch.as_tensor(kinked, device=tg_sub[j].device)
        tg_sub[j][kt] = sub[j].color[kt]
        adj[0][kinked] = 0.0
        g = oracle.adjoint(s, c64, t, adj, grads=("volume", "tf"), method=oracle.EXACT_REPLAY)
        want[k]["v_clean"] = want[k]["v_clean"] + g["volume"][0]
        want[k]["t_clean"] = want[k]["t_clean"] + g["tf"][0]
        sample.append((vol, tf, vp, w, h, steps_w[i], wl.fov_y, tg))
    p.run()
    torch.cuda.synchronize()
    inner = (slice(2, -2),) * 3

    def err(pairs):
        num = sum(float(np.sum((a - b) ** 2)) for a, b in pairs)
        den = sum(float(np.sum(b ** 2)) for _, b in pairs)
        return math.sqrt(num / den)

    acc = {"views": [int(i) for i in idx], "forward_rel_err_vs_fp64": ferr,
           "density_grad_rel_err_vs_fp64": err([(g_v[k].double().cpu().numpy()[inner], want[k]["v_clean"][inner])
                                                for k in used]),
           "tf_grad_rel_err_vs_fp64": err([(g_t[k].double().cpu().numpy(), want[k]["t_clean"]) for k in used]),
           "density_grad_rel_err_all_rays": err([(got_all[k][0][inner], want[k]["v"][inner]) for k in used]),
           "tf_grad_rel_err_all_rays": err([(got_all[k][1], want[k]["t"]) for k in used]),
           "fp32_cpu_twin_density_grad_rel_err_all_rays": err([(want[k]["v_twin"][inner], want[k]["v"][inner])
                                                               for k in used]),
           "kink_samples": kink_samples, "samples": samples,
           "note": "density gradients on interior voxels; headline gradient errors leave out the rays with a "
                   "sample within FP32 rounding of a control point (loss gradient zeroed on both sides)"}
    return acc, sample


# ---------------------------------------------------------------------------
# forward mode on the C2 scene (pytests/tests/stepsize/train_stepsize.py:150-200,
# pytests/tests/camera/test_camera_optimization.py:90-125)
# ---------------------------------------------------------------------------
def cpu_baseline_fwd(wl, D):
    """The FP32 dual-number restatement of the reference's forward mode on view 0."""
    import numpy as np
    from oracle import oracle

    threads = min(16, os.cpu_count() or 1)
    oracle.set_threads(threads)
    s = dict(W=wl.width, H=wl.height, filter=1, camera=2, tf_mode=1, blend=0, early_out=1 - 1e-5, fov_y=wl.fov_y,
             volume=wl.volume.double().cpu().numpy(), box_min=np.full((1, 3), -0.5), box_size=np.ones((1, 3)),
             step=np.full((1, 1, 1), wl.step), tf=wl.tf.double().cpu().numpy(),
             cam_matrix=wl.viewports[:1].double().cpu().numpy())
    t0 = time.perf_counter()
    _, t, _ = oracle.forward_gradients(s, D, d_step=0, d_start=(1, 2, 3), d_dir=(4, 5, 6), dtype=np.float32)
    dt = time.perf_counter() - t0
    samples = float(t.sum())
    return {"value": samples / dt, "unit": "ray-samples/s", "cores": threads, "kind": "port",
            "sample": f"view 0 of the workload ({wl.width}x{wl.height}, {samples:.3g} samples, {dt:.2f} s): FP32 "
                      f"dual-number forward mode with D={D}, the restatement of renderer_kernels.cuh:135-308"}


def bench_fwd(args):
    import numpy as np
    import torch
    from diffdvr_amd import pyrenderer as pr
    from diffdvr_amd import scenes

    dev = torch.device("cuda:0")
    wl = scenes.workload_c2(device="cpu").to(dev)
    B, H, W = wl.B, wl.height, wl.width
    stream = torch.cuda.current_stream()
    steps = args.steps or 20
    args.steps = steps
    D = 7
    target = torch.empty(B, H, W, 4, device=dev)
    tterm = torch.empty(B, H, W, dtype=torch.int32, device=dev)
    pr.Renderer.render_forward(make_inputs(pr, wl, wl.target_tf), pr.RendererOutputs(target, tterm))
    inp = make_inputs(pr, wl, wl.tf)
    fd = pr.ForwardDifferencesSettings()
    fd.D = D
    fd.d_stepsize = 0
    fd.d_rayStart = pr.int3(1, 2, 3)
    fd.d_rayDir = pr.int3(4, 5, 6)
    color = torch.empty(B, H, W, 4, device=dev)
    term = torch.empty(B, H, W, dtype=torch.int32, device=dev)
    grads = torch.empty(B, H, W, D, 4, device=dev)
    outs = pr.RendererOutputs(color, term)
    adj = pr.AdjointOutputs()
    adj.has_stepsize_derivatives = True
    adj.adj_stepsize = torch.zeros(B, H, W, device=dev)
    adj.has_camera_derivatives = True
    adj.adj_camera_ray_start = torch.zeros(B, H, W, 3, device=dev)
    adj.adj_camera_ray_dir = torch.zeros(B, H, W, 3, device=dev)
    n_all = float(B * H * W * 4)

    def render():
        pr.Renderer.render_forward_gradients(inp, fd, outs, grads)

    def fwd_step():
        render()
        adj.adj_stepsize.zero_()
        adj.adj_camera_ray_start.zero_()
        adj.adj_camera_ray_dir.zero_()
        g = torch.sign(color - target) / n_all
        pr.Renderer.forward_variables_to_gradients(grads, g, fd, adj)

    if args.pmc_probe:
        for _ in range(3):
            render()
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=dev)   # 256 MiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"pmc_probe": True, "copy_bytes": src.numel() * 4,
                          "samples_per_step": float(term.sum().item())}))
        return None
    for _ in range(args.warmup):
        fwd_step()
    torch.cuda.synchronize()
    n_samples = float(term.sum().item())
    ms = timed(fwd_step, steps, stream)
    value = n_samples / (ms * 1e-3)
    kernel_ms = float(np.mean([timed(render, 1, stream) for _ in range(max(3, steps // 2))]))
    fcol = torch.empty_like(color)
    fterm = torch.empty_like(term)
    ms_forward = timed(lambda: pr.Renderer.render_forward(inp, pr.RendererOutputs(fcol, fterm), path="tiled"),
                       steps, stream)
    config = {"workload": "forward mode on the C2 scene: 128^3 blob volume, 64x4 texture TF, 8 views 256x256, "
                          "step 0.25 voxel; derivatives of every pixel w.r.t. its step size, ray start and ray "
                          "direction (D = 7), contracted with the L1 loss gradient",
              "views": B, "image": [H, W], "samples_per_step": n_samples, "D": D}
    extra = {"forward_only_ms": ms_forward, "derivative_cost_vs_forward": kernel_ms / ms_forward}
    result = result_line(args, "fwd", value, ms, n_samples, config, kernel_ms, extra,
                         kernel="dvr::k_forward_gradients",
                         metric="forward-mode ray-samples/s on 1 MI355X (D = 7 derivatives per pixel)")
    result["roofline"]["traffic"] = traffic_from_profiles("pmc_traffic_fwdgrad_C2.txt")
    if not args.profile and not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline_fwd(wl, D)
    return result


# ---------------------------------------------------------------------------
# pre-shading step of the density-from-colour reconstruction
# (pytests/tests/volume/preshaded_to_density.py:110-140): colour = TF(density),
# weighted MSE against a target colour volume, backward to the density
# ---------------------------------------------------------------------------
PRESHADE_BYTES = {"forward": 20, "adjoint": 24}   # per voxel: 4 B density (+16 B adj colour) in, 16 B colour / 4 B out


def graph_of(fn, reps):
    """fn called reps times, captured as one HIP graph (fn launches no memset)."""
    import torch
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        fn()
    torch.cuda.current_stream().wait_stream(side)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        for _ in range(reps):
            fn()
    torch.cuda.synchronize()
    return g


def cpu_baseline_preshade(dens_cpu, tf_cpu):
    import numpy as np
    from oracle import oracle

    threads = min(16, os.cpu_count() or 1)
    oracle.set_threads(threads)
    d = dens_cpu[0].numpy().astype(np.float32)
    tf = tf_cpu[0].numpy().astype(np.float32)
    adj = np.ones((4,) + d.shape, dtype=np.float32)
    t0 = time.perf_counter()
    oracle.preshade(3, tf, d, dtype=np.float32)
    oracle.preshade_adjoint(3, tf, d, adj, dtype=np.float32)
    dt = time.perf_counter() - t0
    return {"value": d.size / dt, "unit": "voxels/s", "cores": threads, "kind": "port",
            "sample": f"the whole {d.shape[0]}^3 volume ({dt:.2f} s): FP32 pre-shading and its adjoint with the "
                      f"Gaussian TF, the restatement of renderer/tf_utils_impl.cu:17-121 (OpenMP)"}


def bench_preshade(args):
    import numpy as np
    import torch
    import torch.nn.functional as F
    from diffdvr_amd import pyrenderer as pr
    from diffdvr_amd import scenes
    from diffdvr_amd._native import native as n

    dev = torch.device("cuda:0")
    N = 256
    steps = args.steps or 20
    args.steps = steps
    dens_cpu = scenes.blob_volume(N, 12, seed=41)
    tf_cpu = scenes.gaussian_tf(4, seed=42)
    dens = dens_cpu.to(dev)
    tf = tf_cpu.to(dev)
    target = pr.TFUtils.preshade_volume(scenes.blob_volume(N, 12, seed=43).to(dev), tf, pr.TFMode.Gaussian)
    weight = 0.5 + 0.5 * torch.clamp(target[3:4], 0, 1)
    param = dens.clone().requires_grad_(True)
    stream = torch.cuda.current_stream()
    color = torch.empty(4, N, N, N, device=dev)
    adj_color = torch.rand(4, N, N, N, device=dev)
    adj_dens = torch.empty_like(dens)

    def fwd():
        n.preshade_volume(dens, tf, 3, color)

    def adj():
        n.preshade_volume_adjoint(dens, tf, 3, adj_color, adj_dens)

    def train_step():
        param.grad = None
        c = pr.TFUtils.preshade_volume(param, tf, pr.TFMode.Gaussian)
        loss = torch.mean(F.mse_loss(c, target, reduction="none") * weight)
        loss.backward()

    if args.profile:  # 3 launches of each kernel and one 256 MiB device copy (counter calibration)
        for _ in range(3):
            fwd()
            adj()
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"profile": True, "voxels": N ** 3, "copy_bytes": src.numel() * 4}))
        return None
    for _ in range(args.warmup):
        fwd()
        adj()
        train_step()
    torch.cuda.synchronize()
    # kernel durations: 10 launches captured in a HIP graph, replayed back to back (an eager
    # Python loop leaves the GPU idle between launches of a 60 us kernel)
    reps = 10
    fwd_ms = timed(graph_of(fwd, reps).replay, steps, stream) / reps
    adj_ms = timed(graph_of(adj, reps).replay, steps, stream) / reps
    eager_ms = timed(lambda: (fwd(), adj()), steps, stream)
    step_ms = timed(train_step, steps, stream)
    voxels = float(N ** 3)
    # best fit at the reference script's default of 256 candidates, on 128^3
    c128 = pr.TFUtils.preshade_volume(scenes.blob_volume(128, 12, seed=44).to(dev), tf, pr.TFMode.Gaussian)
    fit_ms = timed(lambda: pr.TFUtils.find_best_fit(c128, tf, pr.TFMode.Gaussian, 256, 1.0), 3, stream)
    # elementwise log-MSE over the colour volume
    lx, ly = torch.log(color + 1e-3), torch.log(target + 1e-3)
    lmse_out = torch.empty_like(lx)
    lmse_ms = timed(graph_of(lambda: n.log_loss(False, lx, ly, lmse_out), reps).replay, steps, stream) / reps
    # importance sampling of 2^22 positions
    locs = torch.rand(1 << 22, 3, device=dev)
    imp_ms = timed(lambda: pr.sample_importance(locs, dens, tf, pr.TFMode.Gaussian, 1, 1, 1, 1, 7), 5, stream)
    achieved = voxels * PRESHADE_BYTES["forward"] / (fwd_ms * 1e-3)
    value = voxels / ((fwd_ms + adj_ms) * 1e-3)
    config = {"workload": "pre-shading of a 256^3 blob density volume through a 4-Gaussian TF and its density "
                          "adjoint (the loss step of preshaded_to_density.py); extras: the full torch loss step, "
                          "best fit (256 candidates, 128^3), log-MSE over 4x256^3, importance sampling of 2^22 "
                          "positions",
              "voxels": N ** 3, "tf": "Gaussian, R = 4"}
    r = {"metric": "pre-shading forward+adjoint voxels/s on 1 MI355X", "value": value, "unit": "voxels/s",
         "n_gpus": 1, "steps": steps, "warmup": args.warmup, "ms_per_step": fwd_ms + adj_ms,
         "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
         "config": config,
         "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                      "frac": achieved / HBM_PEAK, "traffic": traffic_from_profiles("pmc_traffic_preshade_256.txt"),
                      "kernel": "dvr::k_preshade<3, false>", "kernel_ms": fwd_ms,
                      "bytes_per_voxel": PRESHADE_BYTES["forward"]},
         "adjoint_ms": adj_ms, "eager_python_fwd_adj_ms": eager_ms,
         "adjoint_GBps": voxels * PRESHADE_BYTES["adjoint"] / (adj_ms * 1e-3) / 1e9,
         "torch_loss_step_ms": step_ms,
         "best_fit_candidates_per_s": 128 ** 3 * 256 / (fit_ms * 1e-3),
         "log_mse_GBps": lx.numel() * 12 / (lmse_ms * 1e-3) / 1e9,
         "importance_samples_per_s": (1 << 22) / (imp_ms * 1e-3)}
    if not args.no_cpu_baseline:
        r["cpu_baseline"] = cpu_baseline_preshade(dens_cpu, tf_cpu)
    return r


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: only --gpus 1 is supported; the benchmark measures one MI355X "
              f"(sharding the views over several GPUs is designed, not implemented)", file=sys.stderr)
        sys.exit(2)
    fn = {"c1": bench_c1, "c2": bench_c2, "c3": bench_c3, "c4": bench_c4, "c5": bench_c5, "fwd": bench_fwd,
          "preshade": bench_preshade}[args.workload]
    result = fn(args)
    if result is not None:
        print(json.dumps(result))


if __name__ == "__main__":
    main()
