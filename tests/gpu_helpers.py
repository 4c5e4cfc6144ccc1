"""Helpers of the GPU parity tests: an oracle scene dict as pyrenderer inputs on the device."""
from __future__ import annotations

import numpy as np
import torch

from diffdvr_amd import pyrenderer as pr


def dev():
    return torch.device("cuda:0")


def t32(a, device=None):
    return torch.as_tensor(np.asarray(a), dtype=torch.float32, device=device or dev())


def gpu_inputs(s, box_as_tensor=False, step_as_tensor=None, volume_view=None):
    """pyrenderer.RendererInputs of an oracle scene dict (arrays copied to float32 on cuda:0)."""
    d = dev()
    inp = pr.RendererInputs()
    inp.screen_size = pr.int2(int(s["W"]), int(s["H"]))
    inp.volume = volume_view if volume_view is not None else t32(s["volume"], d)
    inp.volume_filter_mode = pr.VolumeFilterMode.Trilinear if s["filter"] == 1 else pr.VolumeFilterMode.Nearest
    if box_as_tensor or s["box_min"].shape[0] > 1:
        inp.box_min = t32(s["box_min"], d)
        inp.box_size = t32(s["box_size"], d)
    else:
        inp.box_min = pr.real3(*[float(v) for v in s["box_min"][0]])
        inp.box_size = pr.real3(*[float(v) for v in s["box_size"][0]])
    cam = int(s["camera"])
    inp.camera_mode = [pr.CameraMode.RayStartDir, pr.CameraMode.InverseViewMatrix, pr.CameraMode.ReferenceFrame][cam]
    if cam == 0:
        inp.camera = pr.CameraPerPixelRays(t32(s["ray_start"], d), t32(s["ray_dir"], d))
    elif cam == 1:
        inp.camera = t32(s["cam_matrix"], d)
    else:
        inp.camera = pr.CameraReferenceFrame(t32(s["cam_matrix"], d), float(s["fov_y"]))
    st = np.asarray(s["step"])
    use_tensor = step_as_tensor if step_as_tensor is not None else st.size > 1
    inp.step_size = t32(st, d) if use_tensor else float(st.reshape(-1)[0])
    inp.tf_mode = [pr.TFMode.Identity, pr.TFMode.Texture, pr.TFMode.Linear, pr.TFMode.Gaussian][int(s["tf_mode"])]
    inp.tf = t32(s["tf"], d)
    inp.blend_mode = pr.BlendMode.BeerLambert if s["blend"] == 0 else pr.BlendMode.Alpha
    inp.blending_early_out = float(s.get("early_out", 1 - 1e-5))
    return inp


def batch(s):
    from oracle import oracle
    return oracle.batch_size(s)


def render(s, path="auto", **kw):
    B, H, W = batch(s), int(s["H"]), int(s["W"])
    inp = gpu_inputs(s, **kw)
    color = torch.full((B, H, W, 4), float("nan"), device=dev())
    term = torch.full((B, H, W), -7, dtype=torch.int32, device=dev())
    pr.Renderer.render_forward(inp, pr.RendererOutputs(color, term), path=path)
    torch.cuda.synchronize()
    return inp, color, term


def rel_err(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-30))


def norm_err(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))
