"""GPU forward parity against the FP64 oracle (renderer_kernels.cuh:22-127).

Protocol: the termination index must agree with the free-running FP64 oracle
for >= 99.9% of the rays (smooth workloads; >= 99% on white-noise volumes)
and by at most one sample everywhere; colours are compared with the oracle
replaying the GPU's sample count (FP64), error = max|gpu - oracle| / max|oracle|:
1e-5 on the C1 workload and on smooth volumes in every camera / filter / TF /
blend mode (tiled path; 4e-5 for the plain path's FP32 formulation), 1e-4 on
white-noise stress scenes (see
tests/test_oracle_render.py::test_fp32_twin_matches_fp64_forward).
"""
import numpy as np
import pytest
import torch

from helpers import oracle, scenes, scene_from_workload, camera_scene
from gpu_helpers import render, rel_err, gpu_inputs, t32, dev
from diffdvr_amd import pyrenderer as pr

pytestmark = pytest.mark.gpu


def check_forward(s, color, term, tol, term_frac):
    t_gpu = term.cpu().numpy()
    c_gpu = color.double().cpu().numpy()
    c_free, t_free = oracle.forward(s)
    assert np.isfinite(c_gpu).all()
    assert np.abs(t_gpu - t_free).max() <= 1
    assert np.mean(t_gpu == t_free) >= term_frac
    c_ora, _ = oracle.forward(s, frozen_term=t_gpu)
    err = rel_err(c_gpu, c_ora)
    assert err < tol, err
    return err


@pytest.mark.parametrize("path", ["tiled", "plain"])
@pytest.mark.parametrize("workload", ["C1", "C2"])
def test_config_forward_parity(path, workload):
    if workload == "C1":
        s = scene_from_workload(scenes.workload_c1())
    else:
        s = scene_from_workload(scenes.workload_c2(views=2, size=96))
    _, color, term = render(s, path=path)
    # the plain path evaluates the reference's own FP32 formulation; the tiled
    # path the transmittance-tracking one (closer to exact arithmetic)
    tol = 1e-5 if path == "tiled" else 4e-5
    err = check_forward(s, color, term, tol, 0.999)
    c32, t32 = oracle.forward(s, dtype=np.float32, frozen_term=term.cpu().numpy())
    err32 = rel_err(color.double().cpu().numpy(), c32)
    print(f"{workload} {path}: forward rel err vs FP64 {err:.2e}, vs FP32 reference twin {err32:.2e}")


@pytest.mark.parametrize("filter", [0, 1])
@pytest.mark.parametrize("tf_mode", [0, 1, 2, 3])
@pytest.mark.parametrize("blend", [0, 1])
@pytest.mark.parametrize("camera", [0, 1, 2])
def test_mode_matrix(filter, tf_mode, blend, camera):
    rng = np.random.default_rng(1000 + 100 * filter + 10 * tf_mode + 3 * blend + camera)
    s = camera_scene(rng, 21, 19, B=2, filter=filter, tf_mode=tf_mode, blend=blend, camera=camera)
    for path in ("tiled", "plain"):
        _, color, term = render(s, path=path)
        check_forward(s, color, term, 1e-4, 0.99)


@pytest.mark.parametrize("filter", [0, 1])
@pytest.mark.parametrize("tf_mode", [0, 1, 2, 3])
@pytest.mark.parametrize("blend", [0, 1])
@pytest.mark.parametrize("camera", [0, 1, 2])
def test_mode_matrix_smooth(filter, tf_mode, blend, camera):
    """Every camera / filter / TF / blend mode on a smooth (blob) volume: the tiled path within
    1e-5 of the FP64 oracle, the plain path (the reference's FP32 formulation) within 4e-5.
    Nearest filtering jumps in value where a sample rounds to the other voxel: rays with a
    sample within FP32 rounding of such an edge (the oracle's kink census) are left out of the
    bound, and at most 0.1% of the samples may be such samples.  At most one ray in a thousand
    (and at least one) may end one sample earlier or later than the free-running oracle."""
    rng = np.random.default_rng(5000 + 100 * filter + 10 * tf_mode + 3 * blend + camera)
    s = camera_scene(rng, 23, 19, B=2, filter=filter, tf_mode=tf_mode, blend=blend, camera=camera, smooth=True,
                     X=20, Y=18, Z=22, step=0.02)
    for path, tol in (("tiled", 1e-5), ("plain", 4e-5)):
        _, color, term = render(s, path=path)
        t_gpu = term.cpu().numpy()
        c_free, t_free = oracle.forward(s)
        c_ora, _ = oracle.forward(s, frozen_term=t_gpu)
        # A ray may stop one sample apart where its transmittance ties with the early-out threshold
        # within rounding; any other disagreement is limited to 0.1% (at least 1) ray.  The tiled path
        # multiplies the transmittance by (1 - alpha) per sample: an FP32 rounding of a near-opaque
        # sample's alpha (0.9-0.99 with the x6 Gaussian TFs here) is amplified alpha / (1 - alpha)
        # times, so T near the threshold carries ~1e-4..1e-3 relative error (measured up to 1.3e-3
        # against FP64); the plain path accumulates the opacity A in FP32 as the reference does,
        # which near 1 is known to ~2^-24 absolute per sample
        thr = 1e-5
        n = np.maximum(t_gpu, t_free).astype(np.float64)
        win = 5e-3 * thr if path == "tiled" else np.maximum(5e-3 * thr, n * 2.0 ** -24)
        tie = (np.abs((1 - c_free[..., 3]) - thr) <= win) | (np.abs((1 - c_ora[..., 3]) - thr) <= win)
        off = t_gpu != t_free
        bad = off & ~tie
        assert np.abs(t_gpu - t_free).max() <= 1 and np.sum(bad) <= max(1, 1e-3 * t_gpu.size), \
            (path, int(off.sum()), int(bad.sum()), [(tuple(int(v) for v in i), int(t_gpu[tuple(i)]),
                                                    int(t_free[tuple(i)]), float(1 - c_free[tuple(i)][3]))
                                                   for i in np.argwhere(bad)[:6]])
        census = oracle.kink_census(s, t_gpu) if filter == 0 else np.zeros_like(t_gpu)
        clean = census == 0
        assert census.sum() <= 1e-3 * t_gpu.sum(), (int(census.sum()), int(t_gpu.sum()))
        err = np.abs(color.double().cpu().numpy() - c_ora)[clean].max() / np.abs(c_ora).max()
        print(f"{path}: smooth forward err {err:.2e}, rays left out {int((~clean).sum())}")
        assert err <= tol, (path, err)


def test_strided_and_broadcast_inputs():
    rng = np.random.default_rng(77)
    s = camera_scene(rng, 20, 18, B=3, tf_mode=2, camera=2)
    s["tf"] = np.concatenate([s["tf"], s["tf"] * 0.9, s["tf"] * 1.1], axis=0)
    s["tf"][..., 4] = s["tf"][0:1, :, 4]  # keep the control-point positions sorted
    s["box_min"] = np.repeat(s["box_min"], 3, axis=0) + rng.uniform(-0.05, 0.05, (3, 3))
    s["box_size"] = np.repeat(s["box_size"], 3, axis=0)
    s["step"] = (0.02 + rng.uniform(0, 0.02, (1, 18, 20))).astype(np.float32).astype(np.float64)
    # volume passed as a non-contiguous view (z-major storage, permuted)
    vol = t32(np.ascontiguousarray(np.transpose(s["volume"], (0, 3, 2, 1))))
    view = vol.permute(0, 3, 2, 1)
    assert not view.is_contiguous()
    for path in ("tiled", "plain"):
        _, color, term = render(s, path=path, box_as_tensor=True, step_as_tensor=True, volume_view=view)
        check_forward(s, color, term, 1e-4, 0.99)


def test_rays_missing_the_box_and_empty_image():
    rng = np.random.default_rng(5)
    s = camera_scene(rng, 9, 7, tf_mode=1, camera=0)
    s["ray_dir"] = -s["ray_dir"]  # all rays point away from the box
    for path in ("tiled", "plain"):
        _, color, term = render(s, path=path)
        assert torch.all(color == 0) and torch.all(term == 0)
    s = camera_scene(rng, 0, 5, tf_mode=1, camera=2)
    inp = gpu_inputs(s)
    color = torch.empty(1, 5, 0, 4, device=dev())
    term = torch.empty(1, 5, 0, dtype=torch.int32, device=dev())
    pr.Renderer.render_forward(inp, pr.RendererOutputs(color, term))
    torch.cuda.synchronize()


def test_validation_errors_raise():
    rng = np.random.default_rng(6)
    s = camera_scene(rng, 8, 8, tf_mode=1, camera=2)
    inp = gpu_inputs(s)
    color = torch.empty(1, 8, 8, 4, device=dev())
    term = torch.empty(1, 8, 8, dtype=torch.int32, device=dev())
    bad = torch.empty(1, 8, 7, 4, device=dev())
    with pytest.raises(RuntimeError, match="color must have 8 entries at dimension 2"):
        pr.Renderer.render_forward(inp, pr.RendererOutputs(bad, term))
    inp.tf = t32(np.zeros((1, 4, 5)))
    with pytest.raises(RuntimeError, match="tf must have 4 entries"):
        pr.Renderer.render_forward(inp, pr.RendererOutputs(color, term))
    inp = gpu_inputs(s)
    inp.volume = inp.volume.cpu()
    with pytest.raises(RuntimeError, match="GPU only"):
        pr.Renderer.render_forward(inp, pr.RendererOutputs(color, term))
    inp = gpu_inputs(s)
    inp.volume = inp.volume.double()
    with pytest.raises(RuntimeError, match="volume must be of type torch.float32"):
        pr.Renderer.render_forward(inp, pr.RendererOutputs(color, term))
