"""Shared test helpers: scenes as oracle dicts, and the same scene as pyrenderer inputs."""
from __future__ import annotations

import math
import os
import sys

import numpy as np
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import oracle  # noqa: E402
from diffdvr_amd import scenes  # noqa: E402


def scene_from_workload(wl, views=None, tf=None, volume=None):
    """Oracle scene dict (float64 numpy) of a scenes.Workload rendered in ReferenceFrame mode."""
    vp = wl.viewports.detach().cpu().double().numpy()
    if views is not None:
        vp = vp[views]
    return dict(
        W=wl.width, H=wl.height, filter=wl.filter, camera=2, tf_mode=wl.tf_mode, blend=wl.blend,
        early_out=1.0 - 1e-5, fov_y=wl.fov_y,
        volume=(wl.volume if volume is None else volume).detach().cpu().double().numpy(),
        box_min=np.full((1, 3), -0.5), box_size=np.ones((1, 3)),
        step=np.full((1, 1, 1), wl.step), tf=(wl.tf if tf is None else tf).detach().cpu().double().numpy(),
        cam_matrix=vp.astype(np.float32).astype(np.float64),
    )


def random_volume(rng: np.random.Generator, X, Y, Z, B=1):
    return rng.uniform(0.0, 1.0, size=(B, X, Y, Z))


def to_f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def camera_scene(rng, W, H, B=1, filter=1, tf_mode=1, blend=0, camera=0, X=12, Y=10, Z=14, R=8,
                 step=0.03, density_scale=1.0, smooth=False):
    """Small random scene whose values are exactly representable in float32.

    smooth=False: white-noise volume (stress case); smooth=True: a few Gaussian blobs.
    """
    if smooth:
        vol = to_f32(scenes.blob_volume((X, Y, Z), 6, int(rng.integers(1 << 30))).double().numpy() * density_scale)
    else:
        vol = to_f32(random_volume(rng, X, Y, Z) * density_scale)
    box_min = np.array([[-0.5, -0.45, -0.55]])
    box_size = np.array([[1.0, 0.9, 1.1]])
    if tf_mode == 0:
        tf = to_f32(np.array([[[rng.uniform(1, 6), rng.uniform(0.5, 2)]]]))
    elif tf_mode == 1:
        tf = to_f32(rng.uniform(0.05, 1.0, size=(1, R, 4)) * np.array([1, 1, 1, 6.0]))
    elif tf_mode == 2:
        pos = np.concatenate([[0.0], np.sort(rng.uniform(0.1, 0.9, R - 2)), [1.0]])
        rgba = rng.uniform(0.05, 1.0, size=(R, 4)) * np.array([1, 1, 1, 6.0])
        tf = to_f32(np.concatenate([rgba, pos[:, None]], axis=1)[None])
    else:
        tf = to_f32(np.concatenate([rng.uniform(0.1, 1.0, size=(R, 4)) * np.array([1, 1, 1, 6.0]),
                                    rng.uniform(0, 1, size=(R, 1)), rng.uniform(0.1, 0.3, size=(R, 1))], axis=1)[None])
    s = dict(W=W, H=H, filter=filter, camera=camera, tf_mode=tf_mode, blend=blend, early_out=1.0 - 1e-5,
             volume=vol, box_min=box_min, box_size=box_size, step=np.full((1, 1, 1), step), tf=tf, fov_y=0.0)
    if camera == 0:
        # rays from an eye below the box's min corner towards jittered points near the centre:
        # every direction component is positive, so rays enter through min faces (p = 0), where
        # the trilinear cell is continuous.  An entry sample on a max face (p = size-1) sits on
        # the kink of the clamped last cell, where central differences and the reference's
        # one-sided adjoint disagree by construction.
        eye = np.array([-1.6, -0.7, -0.9])
        yy, xx = np.meshgrid(np.linspace(-0.6, 0.6, H), np.linspace(-0.6, 0.6, W), indexing="ij")
        target = np.stack([xx * 0.3, yy, xx], axis=-1) + rng.uniform(-0.02, 0.02, size=(H, W, 3))
        d = target - eye
        d /= np.linalg.norm(d, axis=-1, keepdims=True)
        s["ray_start"] = to_f32(np.broadcast_to(eye, (B, H, W, 3)).copy() + rng.uniform(-0.01, 0.01, (B, H, W, 3)))
        s["ray_dir"] = to_f32(np.broadcast_to(d, (B, H, W, 3)).copy())
    elif camera == 2:
        lat = rng.uniform(-0.8, 0.8, B)
        lon = rng.uniform(0, 2 * math.pi, B)
        s["cam_matrix"] = to_f32(scenes.sphere_viewports(lat, lon, 2.2).double().numpy())
        s["fov_y"] = math.radians(40.0)
    else:
        from diffdvr_amd import pyrenderer as pr
        m = np.array(pr.Camera.compute_matrix(pr.real3(-1.2, -0.6, 0.8), pr.real3(0, 0, 0), pr.real3(0, 0, 1), 45, W, H))
        s["cam_matrix"] = to_f32(m[None])
    return s


def marschner_lobb(n=32, fM=6.0, alpha=0.25):
    """renderer/volume_implicit.cpp:28-42 sampled as renderer/volume.cpp:369-387."""
    scale = np.float32(2.0) / np.float32(n - 1)
    c = (np.float32(-1.0) + np.arange(n, dtype=np.float32) * scale).astype(np.float64)
    x, y, z = np.meshgrid(c, c, c, indexing="ij")
    r = np.sqrt(x * x + y * y)
    pr_ = np.cos(2 * math.pi * fM * np.cos(math.pi * r / 2))
    v = ((1 - np.sin(math.pi * z / 2)) + alpha * (1 + pr_)) / (2 * (1 + alpha))
    return v.astype(np.float32).astype(np.float64)[None]


def consistency_tfs(rng):
    """The TF configurations of testConsistencyTF.cpp:71-101 (uniform random entries; the linear TF's
    inner positions are i/(R-1) in integer arithmetic, i.e. 0)."""
    out = []
    for mode, R, C in ((0, 1, 2), (1, 2, 4), (1, 4, 4), (1, 8, 4), (2, 4, 5)):
        tf = to_f32(rng.uniform(0, 1, (1, R, C)))
        if mode == 2:
            tf[0, :, 4] = 0.0
            tf[0, -1, 4] = 1.0
        out.append((mode, tf))
    return out


def sym_err(a, b):
    """Symmetric error of testConsistencyTF.cpp:15-41."""
    a, b = np.asarray(a).ravel(), np.asarray(b).ravel()
    return float((np.abs(a - b) / np.maximum(1.0, np.abs(a) + np.abs(b))).max())
