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

``Volume`` (.cvol files, implicit datasets, mipmaps, histogram) and ``TFUtils``
(TF editor points, piecewise / texture tensors) read and build host data through
the C-ABI; ``TFUtils.preshade_volume`` / ``find_best_fit``, ``sample_importance``
and the losses ``mul_log`` / ``log_mse`` / ``log_l1`` run HIP kernels on device
tensors (the log losses also have their backward, which the reference raises on).

There is no CPU rendering path: every render call needs HIP device tensors
and raises ``RuntimeError`` otherwise.
"""
from __future__ import annotations

import enum
import math

import numpy as np
import torch
import torch.nn.functional as F

from ._native import native as _n

# R*C of the largest TF the tiled / checkpointed / batched kernels stage in LDS (include/dvr.h)
MAX_TILED_TF_ENTRIES = int(_n.MAX_TILED_TF_ENTRIES)

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
                       checkpoint_interval: int = 0, checkpoints: torch.Tensor | None = None,
                       densities: torch.Tensor | None = None) -> None:
        """Simple forward rendering without gradients (renderer/renderer.h:159-161).

        With checkpoints, ``densities`` (>= density_buffer_elements(inputs) floats) receives every
        sample's density for a later checkpointed adjoint."""
        _n.render_forward(inputs, outputs, path, checkpoint_interval, checkpoints, densities)

    @staticmethod
    def render_adjoint(inputs: RendererInputs, outputs_from_forward: RendererOutputs, adj_color: torch.Tensor,
                       adj_outputs: AdjointOutputs, *, method: str = "auto", path: str = "auto",
                       checkpoint_interval: int = 0, checkpoints: torch.Tensor | None = None,
                       densities: torch.Tensor | None = None) -> None:
        """Adjoint rendering; gradients are ACCUMULATED into adj_outputs (renderer/renderer.h:264-268).

        method: "auto"/"checkpoint" (alpha checkpoints, segment recompute) or
        "inversion" (the reference's blend inversion, plain launch structure).
        densities: those the forward call with the same checkpoints wrote (read instead of
        re-interpolating the volume).
        """
        _n.render_adjoint(inputs, outputs_from_forward, adj_color, adj_outputs, method, path,
                          checkpoint_interval, checkpoints, densities)

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
    def density_buffer_elements(inputs: RendererInputs) -> int:
        return _n.density_buffer_elements(inputs)

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


# ---------------------------------------------------------------------------
# Volume (renderer/volume.h:42-205, bindings/bindings.cpp:234-304)
# ---------------------------------------------------------------------------
class ImplicitEquation(enum.IntEnum):
    """Volume::ImplicitEquation (bindings/bindings.cpp:234-249)."""
    MarschnerLobb = 0
    Cube = 1
    Sphere = 2
    InverseSphere = 3
    DingDong = 4
    Endrass = 5
    Barth = 6
    Heart = 7
    Kleine = 8
    Cassini = 9
    Steiner = 10
    CrossCap = 11
    Kummer = 12
    Blobbly = 13
    Tube = 14


def _x_fastest(flat: torch.Tensor, X: int, Y: int, Z: int) -> torch.Tensor:
    """1*X*Y*Z view of a dense x-fastest buffer (strides XYZ, 1, X, X*Y; renderer/volume.cpp:21-33)."""
    return flat.as_strided((1, X, Y, Z), (X * Y * Z, 1, X, X * Y))


class Volume:
    """A density volume with mipmap levels, on the CPU and (after copy_to_gpu) on the device.

    Levels are float32 tensors of shape 1*X*Y*Z with x the fastest axis, as the reference's
    Volume blob (the renderer accepts any strides)."""

    def __init__(self, filename: str | None = None):
        self._world = [1.0, 1.0, 1.0]
        self._cpu: list[torch.Tensor | None] = []
        self._gpu: list[torch.Tensor | None] = []
        if filename is not None:
            (X, Y, Z), voxel, _, _ = _n.cvol_read_header(str(filename))
            flat = torch.empty(X * Y * Z, dtype=torch.float32)
            _n.cvol_read(str(filename), flat.data_ptr(), flat.numel())
            self._cpu = [_x_fastest(flat, X, Y, Z)]
            self._gpu = [None]
            self._world = [voxel[0] * X, voxel[1] * Y, voxel[2] * Z]

    @staticmethod
    def _from_level0(t: torch.Tensor) -> "Volume":
        v = Volume()
        v._cpu = [t]
        v._gpu = [None]
        return v

    @property
    def world_size(self) -> real3:
        return real3(*[float(w) for w in self._world])

    def set_world_size(self, x: float, y: float, z: float) -> None:
        self._world = [float(x), float(y), float(z)]

    @property
    def resolution(self) -> int3:
        _, X, Y, Z = self._cpu[0].shape
        return int3(int(X), int(Y), int(Z))

    def create_mipmap_level(self, level: int) -> None:
        """Level l has size max(1, N // (l+1)) per axis, averaged with adaptive average pooling
        (renderer/volume.cpp:359-365, 421-432)."""
        if level < 0:
            raise ValueError(f"level has to be non-zero, but is {level}")
        while len(self._cpu) <= level:
            self._cpu.append(None)
            self._gpu.append(None)
        if self._cpu[level] is not None:
            return
        _, X, Y, Z = self._cpu[0].shape
        size = (max(1, X // (level + 1)), max(1, Y // (level + 1)), max(1, Z // (level + 1)))
        pooled = torch.nn.functional.adaptive_avg_pool3d(self._cpu[0], size)
        flat = torch.empty(size[0] * size[1] * size[2], dtype=torch.float32)
        lvl = _x_fastest(flat, *size)
        lvl.copy_(pooled)
        self._cpu[level] = lvl

    @staticmethod
    def create_implicit(name: ImplicitEquation, resolution: int, **kwargs) -> "Volume":
        """Volume::createImplicitDataset (renderer/volume_implicit.cpp:218-236)."""
        R = int(resolution)
        flat = torch.empty(R * R * R, dtype=torch.float32)
        names = list(kwargs.keys())
        _n.volume_implicit(int(name), R, names, [float(kwargs[k]) for k in names], flat.data_ptr())
        return Volume._from_level0(_x_fastest(flat, R, R, R))

    @staticmethod
    def from_numpy(array) -> "Volume":
        """A volume from a 3D float32 array indexed [x][y][z] (bindings/bindings.cpp:269-289)."""
        a = np.asarray(array)
        if a.dtype != np.float32:
            raise RuntimeError("Incompatible format: expected a real array!")
        if a.ndim != 3:
            raise RuntimeError("Incompatible buffer dimension, expected 3!")
        X, Y, Z = a.shape
        flat = torch.empty(X * Y * Z, dtype=torch.float32)
        _x_fastest(flat, X, Y, Z)[0].copy_(torch.from_numpy(np.ascontiguousarray(a)))
        return Volume._from_level0(_x_fastest(flat, X, Y, Z))

    def copy_to_gpu(self) -> None:
        for i, t in enumerate(self._cpu):
            if t is not None:
                g = torch.empty_strided(t.shape, t.stride(), dtype=t.dtype, device="cuda")
                g.copy_(t)
                self._gpu[i] = g

    def save(self, filename: str, compression: int = 0) -> None:
        """Volume::save: float voxels, uncompressed (renderer/volume.cpp:86-158)."""
        t = self._cpu[0]
        _, X, Y, Z = t.shape
        flat = torch.empty(X * Y * Z, dtype=torch.float32)
        _x_fastest(flat, X, Y, Z).copy_(t)
        _n.cvol_write(str(filename), flat.data_ptr(), [X, Y, Z], list(self._world), int(compression))

    def getDataCpu(self, level: int) -> torch.Tensor:
        if level >= len(self._cpu) or self._cpu[level] is None:
            raise RuntimeError(f"Mipmap Level {level} was not created")
        return self._cpu[level]

    def getDataGpu(self, level: int) -> torch.Tensor:
        if level >= len(self._cpu) or self._cpu[level] is None:
            raise RuntimeError(f"Mipmap Level {level} was not created")
        if self._gpu[level] is None:
            self.copy_to_gpu()
        return self._gpu[level]

    def extract_histogram(self):
        """Volume::extractHistogram (renderer/volume.cpp:287-356): (bins[512], min, max, non-zero count)."""
        t = self._cpu[0][0]
        bins, mn, mx, nnz = _n.volume_histogram(t.data_ptr(), list(t.shape), list(t.stride()))
        return bins, mn, mx, nnz


# ---------------------------------------------------------------------------
# TF builders (renderer/tf_utils.cpp:12-172, bindings/bindings.cpp:560-569)
# ---------------------------------------------------------------------------
class TFPoint:
    """renderer::TFPoint: density position and (r, g, b, opacity)."""

    def __init__(self, pos: float = 0.0, val: real4 | None = None):
        self.pos = float(pos)
        self.val = val if val is not None else real4(0.0, 0.0, 0.0, 0.0)

    def __repr__(self):
        return f"TFPoint(pos={self.pos}, val=({self.val.x}, {self.val.y}, {self.val.z}, {self.val.w}))"


class _PreshadeVolume(torch.autograd.Function):
    """renderer/tf_utils.cpp:173-247: color (4 x X x Y x Z) = TF(density); backward: the density adjoint."""

    @staticmethod
    def forward(ctx, density, tf, tf_mode):
        out = torch.empty((4,) + tuple(density.shape[1:]), dtype=density.dtype, device=density.device)
        _n.preshade_volume(density, tf, int(tf_mode), out)
        ctx.save_for_backward(density, tf)
        ctx.tf_mode = int(tf_mode)
        return out

    @staticmethod
    def backward(ctx, grad_output):
        density, tf = ctx.saved_tensors
        adj = torch.empty_like(density)
        _n.preshade_volume_adjoint(density, tf, ctx.tf_mode, grad_output, adj)
        return adj, None, None


class TFUtils:
    @staticmethod
    def preshade_volume(density: torch.Tensor, tf: torch.Tensor, tfMode) -> torch.Tensor:
        """Pre-shades a density volume (1 x X x Y x Z) through the TF into a colour volume
        (4 x X x Y x Z: r, g, b, opacity).  Differentiable w.r.t. the density
        (renderer/tf_utils.cpp:249-253, bindings/bindings.cpp:570-581).  tfMode: Identity,
        Texture, Linear, Gaussian or GaussianLog."""
        return _PreshadeVolume.apply(density, tf, tfMode)

    @staticmethod
    def find_best_fit(color_volume: torch.Tensor, tf: torch.Tensor, tf_mode, num_samples: int,
                      opacity_weighting: float, previous_densities: torch.Tensor | None = None,
                      neighbor_weighting: float | None = None) -> torch.Tensor:
        """The density (1 x X x Y x Z) whose TF colour best matches each voxel of the colour
        volume among num_samples random densities; with previous_densities, neighbour
        differences are penalised by neighbor_weighting (renderer/tf_utils.cpp:255-332,
        bindings/bindings.cpp:582-633).  The seed comes from torch's CPU generator."""
        if (previous_densities is None) != (neighbor_weighting is None):
            raise TypeError("previous_densities and neighbor_weighting are given together")
        out = torch.zeros((1,) + tuple(color_volume.shape[1:]), dtype=color_volume.dtype,
                          device=color_volume.device)
        seed = int(torch.randint(0, 2 ** 62, (1,), dtype=torch.int64).item())
        _n.find_best_fit(color_volume, tf, int(tf_mode), int(num_samples), float(opacity_weighting),
                         previous_densities, float(neighbor_weighting or 0.0), seed, out)
        return out

    @staticmethod
    def assemble_from_settings(color_values_lab, color_positions, opacity_values, opacity_positions,
                               min_density: float, max_density: float, opacity_scaling: float,
                               purge_zero_opacity_regions: bool) -> list:
        """Merges the editor's colour axis (L*a*b*, real3) and opacity axis into TF points."""
        lab = []
        for c in color_values_lab:
            lab += [float(c.x), float(c.y), float(c.z)] if hasattr(c, "x") else [float(v) for v in c]
        pts = _n.tf_assemble_from_settings(lab, [float(v) for v in color_positions],
                                           [float(v) for v in opacity_values], [float(v) for v in opacity_positions],
                                           float(min_density), float(max_density), float(opacity_scaling),
                                           bool(purge_zero_opacity_regions))
        return [TFPoint(p, real4(*v)) for p, v in pts]

    @staticmethod
    def get_piecewise_tensor(points) -> torch.Tensor:
        """1*N*5 tensor (r, g, b, opacity, position) of a linear TF (renderer/tf_utils.cpp:120-135)."""
        rows = [[p.val.x, p.val.y, p.val.z, p.val.w, p.pos] for p in points]
        return torch.tensor([rows], dtype=torch.float32)

    @staticmethod
    def get_texture_tensor(points, resolution: int) -> torch.Tensor:
        """1*resolution*4 texture sampled at the texel centres (renderer/tf_utils.cpp:137-172)."""
        out = torch.empty(1, int(resolution), 4, dtype=torch.float32)
        vals = []
        for p in points:
            vals += [p.val.x, p.val.y, p.val.z, p.val.w]
        _n.tf_texture_from_points([p.pos for p in points], vals, int(resolution), out.data_ptr())
        return out

    @staticmethod
    def rgb_to_lab(rgb) -> real3:
        return real3(*_n.rgb_to_lab(float(rgb.x), float(rgb.y), float(rgb.z)))

    @staticmethod
    def lab_to_rgb(lab) -> real3:
        return real3(*_n.lab_to_rgb(float(lab.x), float(lab.y), float(lab.z)))


class _MulLog(torch.autograd.Function):
    """renderer/losses.cpp:8-48: t log2 t (0 below FLT_EPSILON)."""

    @staticmethod
    def forward(ctx, t):
        t = t.contiguous()
        out = torch.empty_like(t)
        _n.mul_log(t, out)
        ctx.save_for_backward(t)
        return out

    @staticmethod
    def backward(ctx, grad_output):
        t, = ctx.saved_tensors
        grad = torch.empty_like(t)
        _n.mul_log_backward(t, grad_output.contiguous(), grad)
        return grad


class _LogLoss(torch.autograd.Function):
    """renderer/losses.cpp:50-142: log-MSE / log-L1 of log-space inputs, with both derivatives."""

    @staticmethod
    def forward(ctx, log_x, log_y, l1):
        if log_x.shape != log_y.shape:
            raise RuntimeError("both args must have the same shape")
        if log_x.device != log_y.device:
            raise RuntimeError("both args must be on the same device")
        if log_x.dtype != log_y.dtype:
            raise RuntimeError("both args must have the same dtype")
        log_x, log_y = log_x.contiguous(), log_y.contiguous()
        out = torch.empty_like(log_x)
        _n.log_loss(bool(l1), log_x, log_y, out)
        ctx.save_for_backward(log_x, log_y)
        ctx.l1 = bool(l1)
        return out

    @staticmethod
    def backward(ctx, grad_output):
        log_x, log_y = ctx.saved_tensors
        gx, gy = torch.empty_like(log_x), torch.empty_like(log_y)
        _n.log_loss_backward(ctx.l1, log_x, log_y, grad_output.contiguous(), gx, gy)
        return gx, gy, None


def mul_log(t: torch.Tensor) -> torch.Tensor:
    """t * log2(t) elementwise, 0 where t < FLT_EPSILON; differentiable (renderer/losses.cpp:45-48)."""
    return _MulLog.apply(t)


def log_mse(log_x: torch.Tensor, log_y: torch.Tensor) -> torch.Tensor:
    """log((x - y)^2) from log x and log y, elementwise; differentiable (renderer/losses.cpp:95-98)."""
    return _LogLoss.apply(log_x, log_y, False)


def log_l1(log_x: torch.Tensor, log_y: torch.Tensor) -> torch.Tensor:
    """log|x - y| from log x and log y, elementwise; differentiable (renderer/losses.cpp:139-142)."""
    return _LogLoss.apply(log_x, log_y, True)


def sample_importance(sample_locations: torch.Tensor, density_volume: torch.Tensor, tf: torch.Tensor, tf_mode,
                      weight_uniform: float, weight_density_gradient: float, weight_opacity: float,
                      weight_opacity_gradient: float, seed: int) -> torch.Tensor:
    """Importance sampling of N positions (N x 3 in [0,1]^3) of the volume: a bool tensor (N,)
    marking the samples taken (renderer/losses.cpp:144-221, bindings/bindings.cpp:661-688).
    Each sample picks a category with probability proportional to its weight: uniform (always
    taken), density-gradient norm, opacity, opacity-gradient norm (taken with probability
    value / maximum over the voxels)."""
    out = torch.empty(sample_locations.shape[0] if sample_locations.dim() >= 1 else 0, dtype=torch.bool,
                      device=sample_locations.device)
    maxima = torch.empty(3, dtype=torch.float32, device=density_volume.device)
    _n.sample_importance(sample_locations, density_volume, tf, int(tf_mode),
                         [float(weight_uniform), float(weight_density_gradient), float(weight_opacity),
                          float(weight_opacity_gradient)], int(seed), maxima, out)
    return out


__all__ = [
    "real3", "real4", "int2", "int3", "VolumeFilterMode", "CameraMode", "TFMode", "BlendMode", "Orientation",
    "CameraPerPixelRays", "CameraReferenceFrame", "mat4", "RendererInputs", "RendererOutputs", "AdjointOutputs",
    "ForwardDifferencesSettings", "GpuTimer", "sync", "use_double_precision", "Renderer", "Camera",
    "ImplicitEquation", "Volume", "TFPoint", "TFUtils", "mul_log", "log_mse", "log_l1", "sample_importance",
]
