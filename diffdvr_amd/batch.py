"""Batched persistent rendering: many views of many volumes in one launch.

DiffDVR renders a batch only when every view shares one image size and one
volume shape (renderer/renderer.cpp:48-391); callers with uneven views loop
over ``Renderer.render_forward`` / ``render_adjoint`` per view
(pytests/diffdvr/renderer.py:127-213) and sum the gradients in torch.  Here a
:class:`BatchPlan` binds the buffers of a heterogeneous batch once, and
:meth:`BatchPlan.run` renders every view, forms the L1/L2 image-loss gradient
in registers and runs the checkpointed adjoint in one persistent HIP launch.

A run allocates nothing and copies nothing from the host, so optimisation
steps built from runs and :class:`Adam` steps can be captured into a HIP graph
(:class:`GraphedSteps`).  The plan holds pointers: parameters updated in place
(by :class:`Adam` or torch) are seen by the next run.

Contract: every tensor is a contiguous float32 (termination int32) HIP tensor
of the plan's device and must stay alive as long as the plan; the plan keeps
references to them.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional, Sequence

import torch

from ._native import native as _n

LOSS_NONE, LOSS_L1, LOSS_L2 = 0, 1, 2
ZERO_GRADIENTS = 1
_TF_COLS = {0: 2, 1: 4, 2: 5, 3: 6}


def _ptr(t: Optional[torch.Tensor]) -> int:
    return 0 if t is None else int(t.data_ptr())


def _check(t: Optional[torch.Tensor], name: str, shape, dtype=torch.float32, device=None, optional=False):
    if t is None:
        if optional:
            return
        raise ValueError(f"{name} must be given")
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if t.device.type != "cuda":
        raise RuntimeError(f"{name} is a host tensor: the batched renderer runs on the GPU only")
    if device is not None and t.device != device:
        raise ValueError(f"{name} is on {t.device}, expected {device}")
    if t.dtype != dtype:
        raise ValueError(f"{name} must be of type {dtype}, but is {t.dtype}")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} must have shape {tuple(shape)}, but has {tuple(t.shape)}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")


@dataclass
class BatchVolume:
    """A density volume X x Y x Z in the box [box_min, box_min + box_size]."""
    data: torch.Tensor
    box_min: Sequence[float] = (-0.5, -0.5, -0.5)
    box_size: Sequence[float] = (1.0, 1.0, 1.0)
    grad: Optional[torch.Tensor] = None      # accumulated d loss / d data


@dataclass
class BatchTf:
    """A transfer function table R x C (C = 2/4/5/6 for identity/texture/linear/Gaussian)."""
    data: torch.Tensor
    grad: Optional[torch.Tensor] = None


@dataclass
class BatchView:
    """One camera / image of the batch.

    ``camera`` is a device tensor: a 3 x 3 viewport (rows eye, right, up) for
    ReferenceFrame cameras, or a 4 x 4 inverse view-projection matrix for
    InverseViewMatrix cameras (row-major, as renderer_camera.cuh:12-61).
    """
    volume: int
    tf: int
    width: int
    height: int
    camera: torch.Tensor
    fov_y: float = 0.0
    step: float = 0.5
    camera_mode: int = 2
    color: Optional[torch.Tensor] = None         # H x W x 4 (allocated if None)
    termination: Optional[torch.Tensor] = None   # H x W int32 (allocated if None)
    target: Optional[torch.Tensor] = None        # H x W x 4, loss target
    adj_color: Optional[torch.Tensor] = None     # H x W x 4, given output adjoint (loss None)
    loss_scale: float = 0.0                      # 0 = mean over all loss elements of the batch
    camera_grad: Optional[torch.Tensor] = None   # 3 x 3 viewport gradient (ReferenceFrame)


def _mode(x) -> int:
    return int(x) if not hasattr(x, "value") else int(x.value)


class BatchPlan:
    """A heterogeneous batch bound to the persistent forward + loss + adjoint kernel.

    Parameters
    ----------
    volumes, tfs, views : the batch tables (views index volumes and tfs)
    volume_filter_mode, tf_mode, blend_mode : shared by all views (pyrenderer enums or ints)
    loss : None (adjoint colour given per view), "l1" or "l2" (mean over all loss elements by default)
    loss_out : optional 1-element float32 tensor; the loss value is added to it (cleared by zeroing runs)
    forward_only : render colour and termination only
    checkpoint_interval, wave_width, grid : kernel options (see :meth:`autotune`)
    density_store : the forward march stores every sample's density for the reverse march (default);
        False re-interpolates them instead (no store: grid x samples-per-ray x 64 floats less memory)
    """

    def __init__(self, volumes: Sequence[BatchVolume], tfs: Sequence[BatchTf], views: Sequence[BatchView], *,
                 volume_filter_mode=1, tf_mode=1, blend_mode=0, blending_early_out: float = 1.0 - 1e-5,
                 loss: Optional[str] = None, loss_out: Optional[torch.Tensor] = None, forward_only: bool = False,
                 checkpoint_interval: int = 16, wave_width: int = 8, grid: int = 0, density_store: bool = True):
        if not views:
            raise ValueError("a batch needs at least one view")
        self.volumes, self.tfs, self.views = list(volumes), list(tfs), list(views)
        tfm = _mode(tf_mode)
        if tfm not in _TF_COLS:
            raise RuntimeError(f"tf mode {tfm} is not supported (Identity=0, Texture=1, Linear=2, Gaussian=3)")
        loss_kind = {None: LOSS_NONE, "l1": LOSS_L1, "l2": LOSS_L2, "mse": LOSS_L2}.get(loss)
        if loss_kind is None:
            raise ValueError("loss must be None, 'l1' or 'l2'")
        dev = self.views[0].camera.device
        if dev.type != "cuda":
            raise RuntimeError("the batched renderer runs on the GPU only")
        self.device = dev
        vt = []
        for i, v in enumerate(self.volumes):
            if v.data.dim() != 3:
                raise ValueError(f"volume {i} must be X x Y x Z")
            _check(v.data, f"volume {i}", v.data.shape, device=dev)
            _check(v.grad, f"volume {i} grad", v.data.shape, device=dev, optional=True)
            vt.append((_ptr(v.data), *[int(s) for s in v.data.shape], [float(x) for x in v.box_min],
                       [float(x) for x in v.box_size], _ptr(None if forward_only else v.grad)))
        tt = []
        for i, t in enumerate(self.tfs):
            if t.data.dim() != 2 or t.data.shape[1] != _TF_COLS[tfm]:
                raise ValueError(f"tf {i} must be R x {_TF_COLS[tfm]}")
            _check(t.data, f"tf {i}", t.data.shape, device=dev)
            _check(t.grad, f"tf {i} grad", t.data.shape, device=dev, optional=True)
            tt.append((_ptr(t.data), int(t.data.shape[0]), int(t.data.shape[1]),
                       _ptr(None if forward_only else t.grad)))
        wt = []
        for i, v in enumerate(self.views):
            H, W = int(v.height), int(v.width)
            if v.color is None:
                v.color = torch.empty(H, W, 4, dtype=torch.float32, device=dev)
            if v.termination is None:
                v.termination = torch.empty(H, W, dtype=torch.int32, device=dev)
            cm = _mode(v.camera_mode)
            _check(v.camera, f"view {i} camera", (3, 3) if cm == 2 else (4, 4), device=dev)
            _check(v.color, f"view {i} color", (H, W, 4), device=dev)
            _check(v.termination, f"view {i} termination", (H, W), dtype=torch.int32, device=dev)
            _check(v.target, f"view {i} target", (H, W, 4), device=dev, optional=True)
            _check(v.adj_color, f"view {i} adj_color", (H, W, 4), device=dev, optional=True)
            _check(v.camera_grad, f"view {i} camera_grad", (3, 3), device=dev, optional=True)
            wt.append((int(v.volume), int(v.tf), W, H, cm, _ptr(v.camera), float(v.fov_y), float(v.step),
                       _ptr(v.color), _ptr(v.termination), _ptr(v.target), _ptr(v.adj_color), float(v.loss_scale),
                       _ptr(None if forward_only else v.camera_grad)))
        _check(loss_out, "loss_out", (1,), device=dev, optional=True)
        self.loss_out = loss_out
        self._h = _n._BatchPlan(vt, tt, wt, _mode(volume_filter_mode), tfm, _mode(blend_mode),
                                float(blending_early_out), loss_kind, _ptr(loss_out), bool(forward_only),
                                dev.index if dev.index is not None else torch.cuda.current_device(),
                                int(checkpoint_interval), int(wave_width), int(grid), not density_store)

    def _stream(self) -> int:
        return int(torch.cuda.current_stream(self.device).cuda_stream)

    def run(self, zero_gradients: bool = True) -> None:
        """Render every view, form the loss gradient and accumulate the adjoints (one persistent launch)."""
        self._h.run(ZERO_GRADIENTS if zero_gradients else 0, self._stream())

    def zero_gradients(self) -> None:
        self._h.zero_gradients(self._stream())

    def kernel_ms(self, zero_gradients: bool = True) -> float:
        """One run with HIP events around the persistent kernel alone (on the current stream);
        waits for it and returns that kernel's duration in milliseconds."""
        return float(self._h.time_kernel(ZERO_GRADIENTS if zero_gradients else 0, self._stream()))

    def freeze(self) -> None:
        """Fix the plan's layout (scratch buffers, kernel arguments) for graph capture: afterwards
        set_options may only repeat the current options and autotune raises."""
        self._h.freeze()

    @property
    def options(self) -> dict:
        k, w, g, ri = self._h.options()
        return {"checkpoint_interval": k, "wave_width": w, "grid": g, "density_store": not ri}

    def set_options(self, checkpoint_interval: int = 16, wave_width: int = 8, grid: int = 0,
                    density_store: bool = True) -> None:
        self._h.set_options(int(checkpoint_interval), int(wave_width), int(grid), not density_store)

    def autotune(self, iterations: int = 3) -> dict:
        """Time the candidates on this batch and keep the fastest: checkpoint interval K in
        {4, 8, 16, 32, 64} jointly with the persistent workgroups (1-4 waves per SIMD and full
        residency, per K), then the wave width (tile 4x16, 8x8, 16x4 or 32x2).  Results are cached
        per device and batch signature (in the process and a file, see dvr_batch_plan_autotune).

        Outputs are overwritten and the gradient buffers left zeroed."""
        k, w, g, ri = self._h.autotune(int(iterations), self._stream())
        return {"checkpoint_interval": k, "wave_width": w, "grid": g, "density_store": not ri}

    @property
    def workspace_bytes(self) -> int:
        return int(self._h.workspace_bytes())

    @property
    def chunk_slots(self) -> int:
        """Slots of the per-wave LDS chunk table of the volume gradients (0: none)."""
        return int(self._h.chunk_slots())


class Adam:
    """torch.optim.Adam (no weight decay, no amsgrad) over several tensors in one HIP launch.

    ``params`` is a list of (param, grad) or (param, grad, columns, clamp) entries:
    ``columns`` restricts the update to the first columns of a 2-D parameter
    (e.g. the RGBA columns of a piecewise-linear TF, leaving its control-point
    positions fixed); ``clamp`` = (lo, hi) clamps the parameter after each step.
    """

    def __init__(self, params, lr=1e-3, betas=(0.9, 0.999), eps=1e-8):
        self.lr, self.betas, self.eps = float(lr), (float(betas[0]), float(betas[1])), float(eps)
        self.state = []
        blocks = []
        dev = None
        for e in params:
            p, g = e[0], e[1]
            cols = e[2] if len(e) > 2 else None
            clamp = e[3] if len(e) > 3 else None
            if dev is None:
                dev = p.device
            _check(p, "param", p.shape, device=dev)
            _check(g, "grad", p.shape, device=dev)
            m = torch.zeros_like(p)
            v = torch.zeros_like(p)
            if cols is None:
                rows, c, stride = 1, p.numel(), p.numel()
            else:
                if p.dim() != 2 or not 0 < cols <= p.shape[1]:
                    raise ValueError("columns needs a 2-D parameter with at least that many columns")
                rows, c, stride = p.shape[0], int(cols), p.shape[1]
            lo, hi = (1.0, 0.0) if clamp is None else (float(clamp[0]), float(clamp[1]))
            self.state.append((p, g, m, v))
            blocks.append((_ptr(p), _ptr(g), _ptr(m), _ptr(v), rows, c, stride, lo, hi))
        if dev is None:
            raise ValueError("Adam needs at least one parameter")
        self.device = dev
        self._h = _n._AdamGroup(blocks, dev.index if dev.index is not None else torch.cuda.current_device())

    def step(self) -> None:
        """One step; the step counter lives on the device (captured steps keep counting)."""
        self._h.step(self.lr, self.betas[0], self.betas[1], self.eps,
                     int(torch.cuda.current_stream(self.device).cuda_stream))

    @property
    def step_count(self) -> int:
        """Steps taken so far (synchronous read of the device counter)."""
        return int(self._h.get_step())


class GraphedSteps:
    """``n`` optimisation steps (clear + batched forward/loss/adjoint + Adam) captured as one HIP graph.

    The plan and the optimiser allocate nothing and copy nothing from the host
    during a step, so the graph holds only kernel nodes (no memset nodes: the
    gradients are cleared by a kernel).  Capturing does not execute the steps;
    each :meth:`replay` runs ``n`` steps, and the Adam step counter (on the
    device) advances by ``n``, exactly as ``n`` eager steps would.
    """

    def __init__(self, plan: BatchPlan, adam: Adam, n: int):
        self.plan, self.adam, self.n = plan, adam, int(n)
        plan.freeze()   # the graph holds the plan's scratch pointers: no re-layout from now on
        self.graph = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream(device=plan.device)
        s.wait_stream(torch.cuda.current_stream(plan.device))
        with torch.cuda.stream(s):
            with torch.cuda.graph(self.graph, stream=s):
                for _ in range(self.n):
                    plan.run(zero_gradients=True)
                    adam.step()
        torch.cuda.current_stream(plan.device).wait_stream(s)

    def replay(self) -> None:
        self.graph.replay()
