"""pyrenderer: the Python API of DiffDVR's renderer on MI355X.

Mirrors the pybind11 module of the reference (bindings/bindings.cpp:132-688):
the classes, enums and ``Renderer`` static methods come from the native
``_pyrenderer`` module (C++ over the C-ABI of ``libdvr_hip.so``); the
``Camera`` helpers that the reference writes with torch ops
(renderer/camera.cpp:87-272) are torch code here as well, with ray
generation and its adjoint on the HIP kernels.

Extensions (keyword-only, defaults keep the reference behaviour):
  * ``Renderer.render_forward(..., path=, checkpoint_interval=, checkpoints=)``
  * ``Renderer.render_adjoint(..., method=, path=, checkpoint_interval=, checkpoints=)``
  * ``Renderer.render_fused(...)``: forward + L1/L2 image loss + adjoint in one launch.

There is no CPU rendering path: every render call needs HIP device tensors
and raises ``RuntimeError`` otherwise.
"""
from __future__ import annotations

import math

import numpy as np
import torch
import torch.nn.functional as F

from ._native import native as _n

real3 = _n.real3
real4 = _n.real4
int2 = _n.int2
int3 = _n.int3
VolumeFilterMode = _n.VolumeFilterMode
CameraMode = _n.CameraMode
TFMode = _n.TFMode
BlendMode = _n.BlendMode
Orientation = _n.Orientation
CameraPerPixelRays = _n.CameraPerPixelRays
CameraReferenceFrame = _n.CameraReferenceFrame
mat4 = _n.mat4
RendererInputs = _n.RendererInputs
RendererOutputs = _n.RendererOutputs
AdjointOutputs = _n.AdjointOutputs
ForwardDifferencesSettings = _n.ForwardDifferencesSettings
GpuTimer = _n.GpuTimer
sync = _n.sync
use_double_precision = _n.use_double_precision
cleanup = _n.cleanup
reload_kernels = _n.reload_kernels
set_cuda_sync_mode = _n.set_cuda_sync_mode
set_cuda_debug_mode = _n.set_cuda_debug_mode
set_cuda_cache_dir = _n.set_cuda_cache_dir
disable_cuda_cache = _n.disable_cuda_cache


def _mat4_to_torch(self) -> torch.Tensor:
    return torch.tensor([self.tolist()], dtype=torch.float32)


def _mat4_array(self, dtype=None, copy=None):
    a = np.array(self.tolist(), dtype=np.float32)
    return a if dtype is None else a.astype(dtype)


mat4.to_torch = _mat4_to_torch
mat4.__array__ = _mat4_array


class Renderer:
    """Static rendering entry points (bindings/bindings.cpp:451-511)."""

    @staticmethod
    def render_forward(inputs: RendererInputs, outputs: RendererOutputs, *, path: str = "auto",
                       checkpoint_interval: int = 0, checkpoints: torch.Tensor | None = None) -> None:
        """Simple forward rendering without gradients (renderer/renderer.h:159-161)."""
        _n.render_forward(inputs, outputs, path, checkpoint_interval, checkpoints)

    @staticmethod
    def render_adjoint(inputs: RendererInputs, outputs_from_forward: RendererOutputs, adj_color: torch.Tensor,
                       adj_outputs: AdjointOutputs, *, method: str = "auto", path: str = "auto",
                       checkpoint_interval: int = 0, checkpoints: torch.Tensor | None = None) -> None:
        """Adjoint rendering; gradients are ACCUMULATED into adj_outputs (renderer/renderer.h:264-268).

        method: "auto"/"checkpoint" (alpha checkpoints, segment recompute) or
        "inversion" (the reference's blend inversion, plain launch structure).
        """
        _n.render_adjoint(inputs, outputs_from_forward, adj_color, adj_outputs, method, path,
                          checkpoint_interval, checkpoints)

    @staticmethod
    def render_fused(inputs: RendererInputs, outputs: RendererOutputs, loss: str, target: torch.Tensor,
                     adj_outputs: AdjointOutputs, *, loss_out: torch.Tensor | None = None, scale: float = 0.0,
                     checkpoint_interval: int = 0) -> None:
        """Forward + per-pixel L1/L2 loss against `target` + checkpointed adjoint in one launch.

        scale 0 means the mean over B*H*W*4 elements (torch's L1Loss / MSELoss).
        The loss value is added to `loss_out` (1-element float32) if given.
        """
        _n.render_fused(inputs, outputs, loss, target, adj_outputs, loss_out, scale, checkpoint_interval)

    @staticmethod
    def checkpoint_buffer_elements(inputs: RendererInputs, checkpoint_interval: int) -> int:
        return _n.checkpoint_buffer_elements(inputs, checkpoint_interval)

    @staticmethod
    def render_forward_gradients(inputs: RendererInputs, differences_settings: ForwardDifferencesSettings,
                                 outputs: RendererOutputs, gradients_out: torch.Tensor) -> None:
        """Forward rendering with forward-mode derivatives (renderer/renderer.h:166-196).

        outputs receive what render_forward writes; gradients_out (B*H*W*D*4)
        the derivative of every pixel's RGBA w.r.t. each of the D traced inputs.
        """
        _n.render_forward_gradients(inputs, differences_settings, outputs, gradients_out)

    @staticmethod
    def forward_variables_to_gradients(forward_variables: torch.Tensor, adj_color: torch.Tensor,
                                       differences_settings: ForwardDifferencesSettings,
                                       adj_outputs: AdjointOutputs) -> None:
        """Contracts forward variables with the colour adjoint into adj_outputs
        (renderer/renderer.h:198-215); gradients are ACCUMULATED."""
        _n.forward_variables_to_gradients(forward_variables, adj_color, differences_settings, adj_outputs)

    @staticmethod
    def compare_to_image(color_input: torch.Tensor, gradients_input: torch.Tensor, color_reference: torch.Tensor,
                         difference_out: torch.Tensor, gradients_out: torch.Tensor, reduce: bool) -> None:
        """L2 distance to a reference image with forward derivatives (renderer/renderer.h:220-248)."""
        _n.compare_to_image(color_input, gradients_input, color_reference, difference_out, gradients_out, reduce)


# ---------------------------------------------------------------------------
# Camera (renderer/camera.cpp:47-272)
# ---------------------------------------------------------------------------
_ORIENTATION_UP = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
_ORIENTATION_PERMUTATION = [(2, -1, -3), (-2, 1, 3), (1, 2, 3), (-1, -2, -3), (-3, -1, 2), (3, 1, -2)]
_ORIENTATION_INVERT_YAW = [True, False, False, True, False, True]
_ORIENTATION_INVERT_PITCH = [False, False, False, False, False, False]


class _GenerateRays(torch.autograd.Function):
    """Per-pixel rays from a B*3*3 viewport (renderer/camera.cpp:200-266)."""

    @staticmethod
    def forward(ctx, viewport, fov_y, width, height):
        if viewport.dim() != 3 or viewport.shape[1] != 3 or viewport.shape[2] != 3:
            raise RuntimeError("viewport must be a tensor with 3 dimensions of shape B*3*3")
        B = viewport.shape[0]
        ray_start = torch.empty(B, height, width, 3, dtype=viewport.dtype, device=viewport.device)
        ray_dir = torch.empty_like(ray_start)
        _n.generate_rays(viewport, float(fov_y), int(width), int(height), ray_start, ray_dir)
        ctx.save_for_backward(viewport)
        ctx.fov_y, ctx.width, ctx.height = float(fov_y), int(width), int(height)
        return ray_start, ray_dir

    @staticmethod
    def backward(ctx, grad_start, grad_dir):
        (viewport,) = ctx.saved_tensors
        if grad_start is None:
            grad_start = torch.zeros(viewport.shape[0], ctx.height, ctx.width, 3, dtype=viewport.dtype,
                                     device=viewport.device)
        if grad_dir is None:
            grad_dir = torch.zeros_like(grad_start)
        grad_viewport = torch.empty(viewport.shape[0], 3, 3, dtype=viewport.dtype, device=viewport.device)
        _n.generate_rays_adjoint(viewport, ctx.fov_y, ctx.width, ctx.height, grad_start, grad_dir, grad_viewport)
        return grad_viewport, None, None, None


def _glm_look_at_lh(eye, center, up):
    f = center - eye
    f = f / np.linalg.norm(f)
    s = np.cross(up, f)
    s = s / np.linalg.norm(s)
    u = np.cross(f, s)
    m = np.eye(4, dtype=np.float32)
    m[0, :3], m[1, :3], m[2, :3] = s, u, f
    m[0, 3], m[1, 3], m[2, 3] = -np.dot(s, eye), -np.dot(u, eye), -np.dot(f, eye)
    return m


def _perspective_fov_lh_zo(fov, width, height, z_near, z_far):
    h = np.float32(math.cos(0.5 * fov) / math.sin(0.5 * fov))
    w = np.float32(h * height / width)
    m = np.zeros((4, 4), dtype=np.float32)
    m[0, 0], m[1, 1] = w, h
    m[2, 2] = z_far / (z_far - z_near)
    m[3, 2] = 1.0
    m[2, 3] = -(z_far * z_near) / (z_far - z_near)
    return m


class Camera:
    """Camera helpers (bindings/bindings.cpp:522-551)."""

    @staticmethod
    def compute_matrix(origin, look_at, up, fov_degrees, width, height, near_clip=0.1, far_clip=10.0) -> mat4:
        """Inverse view-projection matrix for CameraMode.InverseViewMatrix; not differentiable
        (renderer/camera.cpp:87-125: lookAtLH, perspectiveFovLH_ZO)."""
        vec = lambda v: np.array([v.x, v.y, v.z] if hasattr(v, "x") else list(v), dtype=np.float32)
        eye, center, upv = vec(origin), vec(look_at), vec(up)
        upv = upv / np.linalg.norm(upv)
        view = _glm_look_at_lh(eye, center, upv)
        proj = _perspective_fov_lh_zo(math.radians(fov_degrees), float(width), float(height), near_clip, far_clip)
        inv = np.linalg.inv((proj @ view).astype(np.float64)).astype(np.float32)
        return mat4(inv)

    @staticmethod
    def generate_rays(viewport: torch.Tensor, fovY: float, screen_width: int, screen_height: int):
        """Differentiable per-pixel rays (ray_start, ray_dir), each B*H*W*3."""
        return _GenerateRays.apply(viewport, fovY, screen_width, screen_height)

    @staticmethod
    def viewport_from_lookat(origin: torch.Tensor, look_at: torch.Tensor, camera_up: torch.Tensor) -> torch.Tensor:
        """B*3*3 viewport (eye, right, up) from origin / target / up (renderer/camera.cpp:127-147)."""
        for t, n in ((origin, "origin"), (look_at, "lookAt"), (camera_up, "cameraUp")):
            if t.dim() != 2 or t.shape[1] != 3:
                raise RuntimeError(f"{n} must be a tensor with 2 dimensions of shape B*3")
        front = F.normalize(look_at - origin, dim=1)
        right = F.normalize(torch.cross(front, camera_up.expand_as(front), dim=1), dim=1)
        up2 = F.normalize(torch.cross(right, front, dim=1), dim=1)
        return torch.stack([origin.expand_as(front), right, up2], dim=1)

    @staticmethod
    def viewport_from_sphere(center: torch.Tensor, yaw_radians: torch.Tensor, pitch_radians: torch.Tensor,
                             distance: torch.Tensor, orientation) -> torch.Tensor:
        """Viewport of a camera on a sphere around `center` (renderer/camera.cpp:149-198).

        center B*3; yaw, pitch, distance B*1 (size-1 batches broadcast).
        """
        o = int(orientation)
        yaw = yaw_radians * (1.0 if _ORIENTATION_INVERT_YAW[o] else -1.0)
        pitch = pitch_radians * (1.0 if _ORIENTATION_INVERT_PITCH[o] else -1.0)
        pos = [torch.cos(pitch) * torch.cos(yaw) * distance,
               torch.sin(pitch) * distance,
               torch.cos(pitch) * torch.sin(yaw) * distance]
        perm = _ORIENTATION_PERMUTATION[o]
        pos2 = [pos[abs(p) - 1] * (1.0 if p > 0 else -1.0) for p in perm]
        B = max(center.shape[0], yaw.shape[0], pitch.shape[0], distance.shape[0])
        origin = torch.cat(pos2, dim=1).expand(B, 3)
        up = torch.tensor([_ORIENTATION_UP[o]], dtype=center.dtype, device=center.device).expand(B, 3)
        c = center.expand(B, 3)
        return Camera.viewport_from_lookat(origin + c, c, up)


__all__ = [
    "real3", "real4", "int2", "int3", "VolumeFilterMode", "CameraMode", "TFMode", "BlendMode", "Orientation",
    "CameraPerPixelRays", "CameraReferenceFrame", "mat4", "RendererInputs", "RendererOutputs", "AdjointOutputs",
    "ForwardDifferencesSettings", "GpuTimer", "sync", "use_double_precision", "Renderer", "Camera",
]
