"""The torch front end (pytests/diffdvr/renderer.py:61-359) on the GPU.

Gradients obtained with loss.backward() through diffdvr_amd.diffdvr.Renderer
are compared with the FP64 oracle chain:
  * TF and volume: oracle adjoint (exact replay, the GPU's sample counts);
  * camera on a sphere (yaw, pitch, distance): oracle ray adjoint per pixel ->
    oracle generateRays adjoint -> float64 torch autograd of viewport_from_sphere.
The reference differentiates neither tmin nor the sample count w.r.t. the
camera (renderer_kernels.cuh:378-412); the oracle chain does the same.
"""
import math

import numpy as np
import pytest
import torch

from helpers import oracle, scenes
from gpu_helpers import dev, norm_err
from diffdvr_amd import pyrenderer as pr
from diffdvr_amd import diffdvr

pytestmark = pytest.mark.gpu


def base_settings(wl):
    s = pr.RendererInputs()
    s.screen_size = pr.int2(wl.width, wl.height)
    s.volume = wl.volume
    s.volume_filter_mode = pr.VolumeFilterMode.Trilinear
    s.box_min = pr.real3(-0.5, -0.5, -0.5)
    s.box_size = pr.real3(1, 1, 1)
    s.step_size = wl.step
    s.tf_mode = pr.TFMode.Gaussian
    s.tf = wl.tf
    s.blend_mode = pr.BlendMode.BeerLambert
    s.camera_mode = pr.CameraMode.ReferenceFrame
    return s


def smooth_c1():
    """C1's volume and camera setup with a sum-of-Gaussians TF: the gradients of a
    piecewise-linear TF jump where a sample's density crosses a control point
    (renderer_tf.cuh:315-400), so a sample lying within float rounding of one
    makes FP32 and FP64 differ by construction (the FP32 CPU twin of the
    reference shows the same figure); the Gaussian TF has no such kinks."""
    wl = scenes.workload_c1().to(dev())
    wl.tf = scenes.gaussian_tf(4, 1102).to(dev())
    wl.tf[..., 3] *= 8.0
    return wl


def oracle_scene(wl, vp, tf, vol):
    return dict(W=wl.width, H=wl.height, filter=1, camera=2, tf_mode=3, blend=0, early_out=1 - 1e-5,
                fov_y=wl.fov_y, volume=vol.double().cpu().numpy(), box_min=np.full((1, 3), -0.5),
                box_size=np.ones((1, 3)), step=np.full((1, 1, 1), wl.step), tf=tf.double().cpu().numpy(),
                cam_matrix=vp.double().cpu().numpy())


@pytest.mark.parametrize("adjoint_method", ["checkpoint", "inversion"])
def test_tf_and_volume_gradients_through_autograd(adjoint_method):
    wl = smooth_c1()
    wl.width = wl.height = 48
    vp = scenes.sphere_viewports([0.4, -0.3], [0.7, 2.9], 2.5).to(dev())
    tf = wl.tf.clone().requires_grad_(True)
    vol = wl.volume.clone().requires_grad_(True)
    r = diffdvr.Renderer(base_settings(wl), optimize_tf=True, optimize_volume=True, adjoint_method=adjoint_method)
    img = r(camera=vp, fov_y_radians=wl.fov_y, tf=tf, volume=vol)
    rng = np.random.default_rng(3)
    w = torch.as_tensor(rng.standard_normal(img.shape), dtype=torch.float32, device=dev())
    (img * w).sum().backward()
    torch.cuda.synchronize()
    # oracle: same sample counts as the forward pass
    s = oracle_scene(wl, vp, wl.tf, wl.volume)
    color = torch.empty_like(img)
    term = torch.empty(img.shape[:3], dtype=torch.int32, device=dev())
    inp = base_settings(wl)
    inp.camera = pr.CameraReferenceFrame(vp, wl.fov_y)
    pr.Renderer.render_forward(inp, pr.RendererOutputs(color, term))
    t = term.cpu().numpy()
    c, _ = oracle.forward(s, frozen_term=t)
    g = oracle.adjoint(s, c, t, w.double().cpu().numpy(), grads=("tf", "volume"), method=oracle.EXACT_REPLAY)
    etf = norm_err(tf.grad.double().cpu().numpy(), g["tf"])
    evol = norm_err(vol.grad.double().cpu().numpy()[:, 2:-2, 2:-2, 2:-2], g["volume"][:, 2:-2, 2:-2, 2:-2])
    print(f"{adjoint_method}: tf grad err {etf:.2e}, interior volume grad err {evol:.2e}")
    assert etf < 1e-4 and evol < 1e-4


def test_camera_on_sphere_chain():
    """yaw / pitch / distance gradients through CameraOnASphere + Renderer on the GPU against the
    FP64 oracle chain built only from oracle pieces: its viewport_from_sphere (camera.cpp:149-198),
    its ReferenceFrame rays, its ray adjoint (exact replay with the GPU's sample counts), its
    generateRays adjoint (camera_impl.cu:302-357) and the complex-step Jacobian of its viewport."""
    wl = smooth_c1()
    wl.width, wl.height = 40, 32
    cam = diffdvr.CameraOnASphere(pr.Orientation.Yp)
    # cameras in the negative octant (yaw in (pi/2, pi), pitch > 0 with Orientation.Yp): every ray that
    # hits the box enters through a min face, away from the kink of the clamped last trilinear cell
    yaw0, pitch0, dist0 = np.array([[2.2], [2.7]]), np.array([[0.35], [0.55]]), np.array([[2.4], [2.7]])
    yaw, pitch, dist = (torch.tensor(v, dtype=torch.float32, device=dev(), requires_grad=True)
                        for v in (yaw0, pitch0, dist0))
    center = torch.zeros(2, 3, device=dev())
    vp = cam(center, yaw, pitch, dist)
    r = diffdvr.Renderer(base_settings(wl), optimize_camera=True)
    img = r(camera=vp, fov_y_radians=wl.fov_y)
    rng = np.random.default_rng(4)
    w = torch.as_tensor(rng.standard_normal(img.shape), dtype=torch.float32, device=dev())
    (img * w).sum().backward()
    torch.cuda.synchronize()
    # the GPU's sample counts (forward of the same viewport)
    inp = base_settings(wl)
    inp.camera = pr.CameraReferenceFrame(vp.detach(), wl.fov_y)
    color = torch.empty_like(img)
    term = torch.empty(img.shape[:3], dtype=torch.int32, device=dev())
    pr.Renderer.render_forward(inp, pr.RendererOutputs(color, term))
    t = term.cpu().numpy()
    # oracle chain in float64, from the float32 parameters the GPU used
    y64, p64, d64 = (v.detach().double().cpu().numpy() for v in (yaw, pitch, dist))
    vp64 = oracle.viewport_from_sphere(np.zeros((2, 3)), y64, p64, d64, pr.Orientation.Yp)
    s = oracle_scene(wl, torch.as_tensor(vp64), wl.tf, wl.volume)
    rs, rd = oracle.generate_rays(vp64, wl.height, wl.width, wl.fov_y)
    s["camera"] = 0
    s["ray_start"], s["ray_dir"] = rs, rd
    _, t_free = oracle.forward(s)
    assert float((t_free == t).mean()) >= 0.999 and np.abs(t_free - t).max() <= 1
    c, _ = oracle.forward(s, frozen_term=t)
    g = oracle.adjoint(s, c, t, w.double().cpu().numpy(), grads=("camera",), method=oracle.EXACT_REPLAY)
    avp = oracle.generate_rays_adjoint(vp64, wl.height, wl.width, wl.fov_y, g["ray_start"], g["ray_dir"])
    J = oracle.viewport_from_sphere_jacobian(np.zeros((2, 3)), y64, p64, d64, pr.Orientation.Yp)
    want = np.sum(J * avp[..., None], axis=(1, 2))   # B x 3: d/d yaw, pitch, distance
    for k, (got, name) in enumerate(((yaw.grad, "yaw"), (pitch.grad, "pitch"), (dist.grad, "dist"))):
        e = norm_err(got.double().cpu().numpy()[:, 0], want[:, k])
        print(f"camera {name}: grad err {e:.2e}")
        assert e < 1e-4, (name, e)


def test_large_tf_falls_back_to_the_plain_path():
    """A 2048 x 4 texture TF exceeds the LDS table of the tiled kernels (MAX_TILED_TF_ENTRIES):
    diffdvr.Renderer then renders with the plain forward and the inversion adjoint, as the
    reference does for any TF size, instead of failing."""
    wl = smooth_c1()
    wl.width = wl.height = 40
    R = 2048
    assert R * 4 > pr.MAX_TILED_TF_ENTRIES
    u = (np.arange(R) + 0.5) / R
    # smooth in the density, so the derivative jumps at texel edges stay small
    tf_np = np.stack([0.5 + 0.4 * np.sin(6 * u), 0.5 + 0.4 * np.cos(5 * u), 0.5 + 0.3 * np.sin(3 * u + 1),
                      4.0 * (0.2 + 0.8 * u * u)], axis=1)
    tf = torch.tensor(tf_np[None], dtype=torch.float32, device=dev(), requires_grad=True)
    vp = scenes.sphere_viewports([0.4], [0.7], 2.5).to(dev())
    s_in = base_settings(wl)
    s_in.tf_mode = pr.TFMode.Texture
    s_in.tf = tf.detach()
    r = diffdvr.Renderer(s_in, optimize_tf=True)
    img = r(camera=vp, fov_y_radians=wl.fov_y, tf=tf)
    rng = np.random.default_rng(6)
    w = torch.as_tensor(rng.standard_normal(img.shape), dtype=torch.float32, device=dev())
    (img * w).sum().backward()
    torch.cuda.synchronize()
    s = oracle_scene(wl, vp, tf.detach(), wl.volume)
    s["tf_mode"] = 1
    inp = base_settings(wl)
    inp.tf_mode = pr.TFMode.Texture
    inp.tf = tf.detach()
    inp.camera = pr.CameraReferenceFrame(vp, wl.fov_y)
    color = torch.empty_like(img)
    term = torch.empty(img.shape[:3], dtype=torch.int32, device=dev())
    pr.Renderer.render_forward(inp, pr.RendererOutputs(color, term), path="plain")
    t = term.cpu().numpy()
    c, _ = oracle.forward(s, frozen_term=t)
    assert np.abs(img.detach().double().cpu().numpy() - c).max() <= 4e-5 * np.abs(c).max()
    g = oracle.adjoint(s, c, t, w.double().cpu().numpy(), grads=("tf",), method=oracle.EXACT_REPLAY)["tf"]
    census = oracle.kink_census(s, t)
    e = norm_err(tf.grad.double().cpu().numpy(), g)
    print(f"2048x4 TF through the plain path: tf grad err {e:.2e}, kinked rays {int((census > 0).sum())}")
    assert e <= (1e-4 if int(census.sum()) == 0 else 1e-3), e
