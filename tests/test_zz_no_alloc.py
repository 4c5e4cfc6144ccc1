"""The allocation-free fused entry point (dvr_render_fused_noalloc) against dvr_render_fused.

Both are called through the C-ABI (pybind -> libdvr_hip.so) on the same
inputs, twice in a row with the second call on a smaller batch: colour and
termination index must be bitwise equal (same kernels), the loss and the TF /
volume gradients equal to 1e-5 (the allocation-free call re-interpolates the
densities instead of reading a density store: the same values, summed by
atomics in another order).  The call must not change the device's free memory,
and inputs whose checkpoints exceed the static scratch must be rejected.
"""
import numpy as np
import pytest
import torch

from helpers import scenes
from gpu_helpers import dev, norm_err
from diffdvr_amd import pyrenderer as pr

pytestmark = pytest.mark.gpu


def c2_inputs(wl, tf):
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


def fused_call(fn, inp, wl, target, K):
    B, H, W = wl.B, wl.height, wl.width
    color = torch.full((B, H, W, 4), float("nan"), device=dev())
    term = torch.full((B, H, W), -1, dtype=torch.int32, device=dev())
    adj = pr.AdjointOutputs()
    adj.has_tf_derivatives = True
    adj.adj_tf = torch.zeros_like(wl.tf)
    adj.has_volume_derivatives = True
    adj.adj_volume = torch.zeros_like(wl.volume)
    loss = torch.zeros(1, device=dev())
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info(dev())[0]
    fn(inp, pr.RendererOutputs(color, term), "l2", target, adj, loss, 0.0, K)
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info(dev())[0]
    return color, term, adj.adj_tf, adj.adj_volume, loss, free0 - free1


def first_call_on_device():
    """The first call on a device looks up the static scratch (hipGetSymbolAddress), which makes the
    runtime load the array's code object: the device then holds the array until the process ends.
    Later calls must not change the device's free memory."""
    wl = scenes.workload_c2(views=1, size=16).to(dev())
    target = torch.zeros(1, 16, 16, 4, device=dev())
    fused_call(pr._n.render_fused_noalloc, c2_inputs(wl, wl.tf), wl, target, 16)


@pytest.mark.parametrize("K", [16, 64])
def test_noalloc_equals_fused_twice_then_smaller(K):
    first_call_on_device()
    for views, size in ((4, 128), (2, 72)):   # the second call: a smaller batch and image
        print(f"case K={K} views={views} size={size}", flush=True)
        wl = scenes.workload_c2(views=views, size=size).to(dev())
        target = torch.empty(wl.B, size, size, 4, device=dev())
        tterm = torch.empty(wl.B, size, size, dtype=torch.int32, device=dev())
        pr.Renderer.render_forward(c2_inputs(wl, wl.target_tf), pr.RendererOutputs(target, tterm))
        inp = c2_inputs(wl, wl.tf)
        need = pr.Renderer.checkpoint_buffer_elements(inp, K)
        assert 0 < need <= pr._n.NOALLOC_SCRATCH_FLOATS
        ref = fused_call(pr._n.render_fused, inp, wl, target, K)
        new = fused_call(pr._n.render_fused_noalloc, inp, wl, target, K)
        assert torch.equal(ref[0], new[0]) and torch.equal(ref[1], new[1])
        assert abs(float(ref[4]) - float(new[4])) <= 1e-6 * abs(float(ref[4]))
        e_tf = norm_err(new[2].double().cpu().numpy(), ref[2].double().cpu().numpy())
        e_vol = norm_err(new[3].double().cpu().numpy(), ref[3].double().cpu().numpy())
        print(f"  TF grad diff {e_tf:.1e}, volume grad diff {e_vol:.1e}, device memory taken by the call {new[5]}")
        assert e_tf <= 1e-5 and e_vol <= 1e-5, (e_tf, e_vol)
        assert new[5] == 0, new[5]


def test_noalloc_rejects_inputs_beyond_its_scratch():
    wl = scenes.workload_c2(views=1, size=2048).to(dev())
    inp = c2_inputs(wl, wl.tf)
    assert pr.Renderer.checkpoint_buffer_elements(inp, 8) > pr._n.NOALLOC_SCRATCH_FLOATS
    target = torch.zeros(1, 2048, 2048, 4, device=dev())
    color = torch.empty_like(target)
    term = torch.empty(1, 2048, 2048, dtype=torch.int32, device=dev())
    adj = pr.AdjointOutputs()
    adj.has_tf_derivatives = True
    adj.adj_tf = torch.zeros_like(wl.tf)
    with pytest.raises(RuntimeError, match="static scratch holds"):
        pr._n.render_fused_noalloc(inp, pr.RendererOutputs(color, term), "l2", target, adj, None, 0.0, 8)
