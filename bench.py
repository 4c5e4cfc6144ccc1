"""Throughput of DiffDVR's differentiable ray marcher on one MI355X.

Metric: forward+adjoint ray-samples/s of the batched, checkpointed, fused
path (one launch: forward march, per-pixel L1 loss, checkpointed reverse march
with an LDS-privatised TF gradient) against the plain per-view form (per view:
reference-structured forward launch, torch loss gradient, reference-structured
inversion adjoint launch).  A ray-sample is one blended march iteration;
N_samples = sum of the termination indices of the forward pass, and the
adjoint processes the same N.

Workload (default, the TF-gradient configuration): 128^3 float32 density of
32 seeded Gaussian blobs, 64x4 texture TF (U[0,1]), 8 cameras at latitude 20
degrees evenly spaced in longitude, radius 2.5, 256x256 images, step 0.25 voxel,
trilinear, Beer-Lambert, L1 loss against an image rendered with a second
seeded TF; gradient w.r.t. the 64x4 TF.  Synthetic data, identical on every host.

Prints ONE JSON line (rank 0).  Options:
  --steps K --warmup W   timed / untimed steps (a step = one pass over the batch)
  --gpus N               only N=1 is supported (one GPU per box)
  --no-cpu-baseline      skip the CPU baseline leg
  --profile              short run for rocprofv3 (no baseline, no plain leg)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12            # MI355X HBM3E, bytes/s
BYTES_PER_SAMPLE_TF = 64     # 32 B forward gathers + 32 B adjoint re-gathers (TF gradient privatised)
METRIC = "fwd+adjoint ray-samples/s on 1 MI355X, batched vs plain, at 1e-4 rel grad err"


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--profile", action="store_true")
    p.add_argument("--checkpoint-interval", type=int, default=0)
    p.add_argument("--pmc-probe", action="store_true",
                   help="for rocprofv3 --pmc: 3 fused steps and one 256 MiB device copy (counter calibration)")
    return p.parse_args()


def make_inputs(pr, wl, tf):
    inp = pr.RendererInputs()
    inp.screen_size = pr.int2(wl.width, wl.height)
    inp.volume = wl.volume
    inp.volume_filter_mode = pr.VolumeFilterMode.Trilinear
    inp.box_min = pr.real3(-0.5, -0.5, -0.5)
    inp.box_size = pr.real3(1.0, 1.0, 1.0)
    inp.camera_mode = pr.CameraMode.ReferenceFrame
    inp.camera = pr.CameraReferenceFrame(wl.viewports, wl.fov_y)
    inp.step_size = wl.step
    inp.tf_mode = pr.TFMode.Texture
    inp.tf = tf
    inp.blend_mode = pr.BlendMode.BeerLambert
    return inp


def traffic_from_profiles(name):
    """HBM bytes per launch of the fused kernel from the committed PMC summary, if present."""
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


def cpu_baseline(wl, target, color, term, adj_tf, scale):
    """The FP32 OpenMP restatement of renderer_cpu.cpp on the whole workload
    (forward + L1 gradient + inversion adjoint with TF gradient), and the FP64
    error of the GPU's fused result (same sample counts)."""
    import numpy as np
    from oracle import oracle

    threads = min(16, os.cpu_count() or 1)
    oracle.set_threads(threads)
    s = dict(W=wl.width, H=wl.height, filter=1, camera=2, tf_mode=1, blend=0, early_out=1 - 1e-5, fov_y=wl.fov_y,
             volume=wl.volume.double().cpu().numpy(), box_min=np.full((1, 3), -0.5), box_size=np.ones((1, 3)),
             step=np.full((1, 1, 1), wl.step), tf=wl.tf.double().cpu().numpy(),
             cam_matrix=wl.viewports.double().cpu().numpy())
    tg = target.double().cpu().numpy()
    # timed: FP32 forward + loss gradient + FP32 inversion adjoint
    t0 = time.perf_counter()
    c32, t32 = oracle.forward(s, dtype=np.float32)
    adj = (np.sign(c32 - tg) * scale).astype(np.float32)
    oracle.adjoint(s, c32, t32, adj, grads=("tf",), dtype=np.float32)
    dt = time.perf_counter() - t0
    samples = float(t32.sum())
    # accuracy of the GPU's fused result (FP64, the GPU's sample counts, the kernel's loss sign)
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


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: only --gpus 1 is supported; this build measures one MI355X "
              f"(multi-GPU sharding is a design note, not a build target)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from diffdvr_amd import pyrenderer as pr
    from diffdvr_amd import scenes

    dev = torch.device("cuda:0")
    wl = scenes.workload_c2(device="cpu").to(dev)
    B, H, W = wl.B, wl.height, wl.width
    K = args.checkpoint_interval
    stream = torch.cuda.current_stream()

    # target images (second seeded TF), outside any timing
    target = torch.empty(B, H, W, 4, device=dev)
    tterm = torch.empty(B, H, W, dtype=torch.int32, device=dev)
    pr.Renderer.render_forward(make_inputs(pr, wl, wl.target_tf), pr.RendererOutputs(target, tterm))

    inp = make_inputs(pr, wl, wl.tf)
    color = torch.empty(B, H, W, 4, device=dev)
    term = torch.empty(B, H, W, dtype=torch.int32, device=dev)
    adj_tf = torch.zeros_like(wl.tf)
    adj = pr.AdjointOutputs()
    adj.has_tf_derivatives = True
    adj.adj_tf = adj_tf
    outs = pr.RendererOutputs(color, term)

    def batched_step():
        adj_tf.zero_()
        pr.Renderer.render_fused(inp, outs, "l1", target, adj, checkpoint_interval=K)

    if args.pmc_probe:
        for _ in range(3):
            batched_step()
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=dev)   # 256 MiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"pmc_probe": True, "copy_bytes": src.numel() * 4,
                          "samples_per_step": float(term.sum().item())}))
        return

    # ---- batched fused path ------------------------------------------------------------------
    for _ in range(args.warmup):
        batched_step()
    torch.cuda.synchronize()
    n_samples = float(term.sum().item())
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record(stream)
    for _ in range(args.steps):
        batched_step()
    ev1.record(stream)
    torch.cuda.synchronize()
    ms_batched = ev0.elapsed_time(ev1) / args.steps
    value = n_samples / (ms_batched * 1e-3)

    # fused kernel alone (one launch per call), HIP events on the launch stream
    k_ms = []
    for _ in range(max(3, args.steps // 2)):
        adj_tf.zero_()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        pr.Renderer.render_fused(inp, outs, "l1", target, adj, checkpoint_interval=K)
        e1.record(stream)
        torch.cuda.synchronize()
        k_ms.append(e0.elapsed_time(e1))
    kernel_ms = float(np.mean(k_ms))
    achieved = n_samples * BYTES_PER_SAMPLE_TF / (kernel_ms * 1e-3)

    result = {
        "metric": METRIC, "value": value, "unit": "ray-samples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_batched, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "C2: 128^3 blob volume, 64x4 texture TF gradient, 8 views 256x256, step 0.25 voxel, "
                               "L1 loss, fused checkpointed fwd+adjoint",
                   "views": B, "image": [H, W], "volume": list(wl.volume.shape[1:]), "samples_per_step": n_samples,
                   "checkpoint_interval": K or 16},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic_from_profiles("pmc_traffic_fused_C2.txt"),
                     "kernel": "dvr::k_adjoint_tiled (fused)", "kernel_ms": kernel_ms,
                     "bytes_per_sample": BYTES_PER_SAMPLE_TF},
    }

    if not args.profile:
        # ---- plain per-view path (reference structure) ----------------------------------------
        views = []
        for b in range(B):
            ib = make_inputs(pr, wl, wl.tf)
            ib.camera = pr.CameraReferenceFrame(wl.viewports[b:b + 1].contiguous(), wl.fov_y)
            cb = torch.empty(1, H, W, 4, device=dev)
            tb = torch.empty(1, H, W, dtype=torch.int32, device=dev)
            views.append((ib, cb, tb, target[b:b + 1].contiguous()))
        padj_tf = torch.zeros_like(wl.tf)
        padj = pr.AdjointOutputs()
        padj.has_tf_derivatives = True
        padj.tf_delayed_accumulation = True   # diffdvr.Renderer default (pytests/diffdvr/renderer.py:270)
        padj.adj_tf = padj_tf
        n_all = float(B * H * W * 4)

        def plain_step():
            padj_tf.zero_()
            for ib, cb, tb, tgb in views:
                o = pr.RendererOutputs(cb, tb)
                pr.Renderer.render_forward(ib, o, path="plain")
                g = torch.sign(cb - tgb) / n_all
                pr.Renderer.render_adjoint(ib, o, g, padj, method="inversion", path="plain")

        for _ in range(max(1, args.warmup)):
            plain_step()
        torch.cuda.synchronize()
        p_steps = max(3, args.steps // 4)
        ev0.record(stream)
        for _ in range(p_steps):
            plain_step()
        ev1.record(stream)
        torch.cuda.synchronize()
        ms_plain = ev0.elapsed_time(ev1) / p_steps
        plain_value = n_samples / (ms_plain * 1e-3)
        plain_err = float(torch.linalg.norm(padj_tf - adj_tf) / torch.linalg.norm(adj_tf))
        result["plain"] = {"value": plain_value, "ms_per_step": ms_plain, "speedup_batched_vs_plain": value / plain_value,
                           "tf_grad_rel_diff_vs_batched": plain_err}

        if not args.no_cpu_baseline:
            # the batched buffers hold the last fused step: colour, termination, TF gradient
            batched_step()
            torch.cuda.synchronize()
            base, ferr, gerr = cpu_baseline(wl, target, color, term, adj_tf, 1.0 / n_all)
            result["cpu_baseline"] = base
            result["accuracy"] = {"forward_rel_err_vs_fp64": ferr, "tf_grad_rel_err_vs_fp64": gerr}
    print(json.dumps(result))


if __name__ == "__main__":
    main()
