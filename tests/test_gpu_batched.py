"""Batched persistent launch (many views of many volumes in one kernel) on the GPU.

Each view of a heterogeneous batch (two volumes of different shapes, two TFs,
views of different image sizes and step sizes) is checked against the FP64
oracle rendering that view alone, the way a reference caller would loop over
Renderer.render_forward / render_adjoint per view and sum the gradients
(pytests/diffdvr/renderer.py:127-213):
  * forward: termination index vs the free-running oracle, colour vs the
    oracle replaying the GPU's sample count, <= 1e-5 relative;
  * adjoint: TF and volume gradients summed over views, and the 3x3 viewport
    gradient of every ReferenceFrame camera (oracle ray adjoint followed by the
    oracle's generateRays adjoint, camera_impl.cu:302-357), normwise <= 1e-4.
Cameras sit in the negative octant so every ray enters the boxes through min
faces, away from the kink of the clamped last trilinear cell.  Also covered:
bitwise agreement with the tiled single-batch kernel, the fused L1/L2 loss,
the autotuner, the multi-block Adam step against torch.optim.Adam, and
captured optimisation steps replayed as a HIP graph.
"""
import math

import numpy as np
import pytest
import torch

from helpers import oracle, scenes, to_f32
from gpu_helpers import dev, norm_err, t32
from diffdvr_amd import batch as bt
from diffdvr_amd import pyrenderer as pr

pytestmark = pytest.mark.gpu

VOL_SHAPES = [(18, 16, 20), (12, 14, 10)]
BOXES = [((-0.5, -0.45, -0.55), (1.0, 0.9, 1.1)), ((-0.4, -0.5, -0.5), (0.8, 1.0, 1.0))]
# (volume, tf, W, H, step)
VIEWS = [(0, 0, 23, 17, 0.02), (1, 1, 16, 21, 0.03), (0, 1, 30, 9, 0.025), (1, 0, 11, 13, 0.02)]
FOV = math.radians(40.0)


def make_tf(rng, tf_mode, R=8, spaced=True):
    """Random TF.  Linear TFs get control points at least 0.5/(R-1) apart unless
    spaced=False: the slope 1/(pos[k+1]-pos[k]) multiplies FP32 density error, so
    nearby control points make any FP32 evaluation ill-conditioned."""
    if tf_mode == 0:
        return to_f32(np.array([[rng.uniform(1, 6), rng.uniform(0.5, 2)]]))
    if tf_mode == 1:
        return to_f32(rng.uniform(0.05, 1.0, size=(R, 4)) * np.array([1, 1, 1, 6.0]))
    if tf_mode == 2:
        if spaced:
            inner = (np.arange(1, R - 1) + rng.uniform(-0.25, 0.25, R - 2)) / (R - 1)
        else:
            inner = np.sort(rng.uniform(0.1, 0.9, R - 2))
        pos = np.concatenate([[0.0], inner, [1.0]])
        rgba = rng.uniform(0.05, 1.0, size=(R, 4)) * np.array([1, 1, 1, 6.0])
        return to_f32(np.concatenate([rgba, pos[:, None]], axis=1))
    return to_f32(np.concatenate([rng.uniform(0.1, 1.0, size=(R, 4)) * np.array([1, 1, 1, 6.0]),
                                  rng.uniform(0, 1, size=(R, 1)), rng.uniform(0.1, 0.3, size=(R, 1))], axis=1))


class Case:
    """A heterogeneous batch as numpy (oracle) and device tensors (plan)."""

    def __init__(self, seed, filter, tf_mode, blend, grads=("tf", "volume", "camera"), loss=None, spaced=True,
                 tf_rows=8):
        rng = np.random.default_rng(seed)
        self.filter, self.tf_mode, self.blend = filter, tf_mode, blend
        self.vols = [to_f32(scenes.blob_volume(sh, 6, int(rng.integers(1 << 30)))[0].double().numpy())
                     for sh in VOL_SHAPES]
        self.tfs = [make_tf(rng, tf_mode, R=tf_rows, spaced=spaced) for _ in range(2)]
        if blend == 1:
            for t in self.tfs:
                t[..., 0 if tf_mode == 0 else 3] *= 0.5
        yaw = rng.uniform(2.2, 2.7, len(VIEWS))
        pitch = rng.uniform(0.35, 0.55, len(VIEWS))
        dist = rng.uniform(2.0, 2.6, len(VIEWS))
        vps = pr.Camera.viewport_from_sphere(torch.zeros(len(VIEWS), 3, dtype=torch.float64),
                                             torch.as_tensor(yaw).reshape(-1, 1), torch.as_tensor(pitch).reshape(-1, 1),
                                             torch.as_tensor(dist).reshape(-1, 1), pr.Orientation.Yp)
        self.vps = to_f32(vps.numpy())
        self.adj_colors = [rng.standard_normal((h, w, 4)) for (_, _, w, h, _) in VIEWS]
        self.targets = [rng.uniform(0, 1, (h, w, 4)) for (_, _, w, h, _) in VIEWS]
        self.grads = grads
        d = dev()
        self.t_vols = [t32(v, d) for v in self.vols]
        self.t_tfs = [t32(t, d) for t in self.tfs]
        self.g_vols = [torch.zeros_like(v) if "volume" in grads else None for v in self.t_vols]
        self.g_tfs = [torch.zeros_like(t) if "tf" in grads else None for t in self.t_tfs]
        self.t_vps = [t32(v, d) for v in self.vps]
        self.g_cams = [torch.zeros(3, 3, device=d) if "camera" in grads else None for _ in VIEWS]
        self.loss = loss
        self.loss_out = torch.zeros(1, device=d)
        views = []
        for i, (vi, ti, w, h, st) in enumerate(VIEWS):
            views.append(bt.BatchView(volume=vi, tf=ti, width=w, height=h, camera=self.t_vps[i], fov_y=FOV, step=st,
                                      adj_color=None if loss else t32(self.adj_colors[i], d),
                                      target=t32(self.targets[i], d) if loss else None,
                                      camera_grad=self.g_cams[i]))
        self.views = views

    def plan(self, forward_only=False, **kw):
        vols = [bt.BatchVolume(v, BOXES[i][0], BOXES[i][1], g) for i, (v, g) in enumerate(zip(self.t_vols, self.g_vols))]
        tfs = [bt.BatchTf(t, g) for t, g in zip(self.t_tfs, self.g_tfs)]
        return bt.BatchPlan(vols, tfs, self.views, volume_filter_mode=self.filter, tf_mode=self.tf_mode,
                            blend_mode=self.blend, loss=self.loss, loss_out=self.loss_out if self.loss else None,
                            forward_only=forward_only, **kw)

    def scene(self, i):
        vi, ti, w, h, st = VIEWS[i]
        return dict(W=w, H=h, filter=self.filter, camera=2, tf_mode=self.tf_mode, blend=self.blend,
                    early_out=1 - 1e-5, fov_y=FOV, volume=self.vols[vi][None], box_min=np.array([BOXES[vi][0]]),
                    box_size=np.array([BOXES[vi][1]]), step=np.full((1, 1, 1), st), tf=self.tfs[ti][None],
                    cam_matrix=self.vps[i][None].astype(np.float64))


def oracle_batch(case, terms, adj_colors):
    """Per-view oracle forward (GPU sample counts) and adjoint, summed like a per-view caller would."""
    g_vol = [np.zeros(v.shape) for v in case.vols]
    g_tf = [np.zeros(t.shape) for t in case.tfs]
    g_cam, colors = [], []
    for i, (vi, ti, w, h, st) in enumerate(VIEWS):
        s = case.scene(i)
        t = terms[i][None]
        c, _ = oracle.forward(s, frozen_term=t)
        colors.append(c[0])
        g = oracle.adjoint(s, c, t, adj_colors[i][None], grads=("tf", "volume", "camera"), method=oracle.EXACT_REPLAY)
        g_vol[vi] += g["volume"][0]
        g_tf[ti] += g["tf"][0]
        g_cam.append(oracle.generate_rays_adjoint(case.vps[i][None].astype(np.float64), h, w, FOV, g["ray_start"],
                                                  g["ray_dir"])[0])
    return colors, g_vol, g_tf, g_cam


def check_against_oracle(case, twin_bound=False):
    """Forward, termination and every requested gradient of a run plan against the per-view oracle.
    twin_bound: the forward bound of a view is max(1e-5, 3x the FP32 CPU twin's error) (for TFs whose
    slope amplifies FP32 density rounding beyond 1e-5 in any FP32 evaluation)."""
    filter = case.filter
    terms = [v.termination.cpu().numpy() for v in case.views]
    colors, g_vol, g_tf, g_cam = oracle_batch(case, terms, case.adj_colors)
    # samples within FP32 rounding of a kink of the reference's derivative (TF clamp /
    # texel edge / control point, alpha clamp): FP32 may take the other side there
    census = [oracle.kink_census(case.scene(i), terms[i][None])[0] for i in range(len(VIEWS))]
    kinks = [int(c.sum()) for c in census]
    print("kink samples per view", kinks)
    for i, v in enumerate(case.views):
        c_free, t_free = oracle.forward(case.scene(i))
        # a ray may stop one sample apart where its transmittance ties with the early-out threshold
        # within FP32 rounding (the window of the smooth mode matrix: the transmittance form carries
        # up to ~1e-3 relative error near a threshold of 1e-5 after near-opaque samples); any other
        # disagreement is limited to 0.1% of the view's rays
        thr, win = 1e-5, 5e-3 * 1e-5
        tie = (np.abs((1 - c_free[0][..., 3]) - thr) <= win) | (np.abs((1 - colors[i][..., 3]) - thr) <= win)
        off = t_free[0] != terms[i]
        bad = off & ~tie
        assert np.abs(t_free[0] - terms[i]).max() <= 1 and bad.sum() <= 1e-3 * terms[i].size, \
            (i, int(off.sum()), int(bad.sum()), [float(1 - c_free[0][tuple(j)][3]) for j in np.argwhere(off)[:4]])
        c = v.color.double().cpu().numpy()
        scale = max(np.abs(colors[i]).max(), 1e-30)
        clean = census[i] == 0  # rays without a sample on a (nearest-filter) value discontinuity
        err = np.abs(c - colors[i])[clean].max() / scale
        err_all = np.abs(c - colors[i]).max() / scale
        bound = 1e-5
        if twin_bound:
            c32, _ = oracle.forward(case.scene(i), dtype=np.float32, frozen_term=terms[i][None])
            twin = np.abs(c32[0].astype(np.float64) - colors[i])[clean].max() / scale
            print(f"view {i}: forward err {err:.2e}, FP32 twin {twin:.2e}")
            bound = max(bound, 3.0 * twin)
        assert err <= bound and err_all <= 1e-3, ("forward", i, err, err_all, bound, int((~clean).sum()))
    errs = {}
    for k, (got, want) in enumerate(zip(case.g_tfs, g_tf)):
        if got is not None:
            errs[f"tf{k}"] = norm_err(got.double().cpu().numpy(), want)
    for k, (got, want) in enumerate(zip(case.g_vols, g_vol)):
        if got is not None:
            errs[f"vol{k}"] = norm_err(got.double().cpu().numpy()[2:-2, 2:-2, 2:-2], want[2:-2, 2:-2, 2:-2])
    for k, (got, want) in enumerate(zip(case.g_cams, g_cam)):
        if got is None:
            continue
        g = got.double().cpu().numpy()
        if filter == 0:  # nearest: no position gradient (renderer_interpolation.cuh:95-106)
            assert np.abs(g).max() == 0.0 and np.abs(want).max() == 0.0
        else:
            errs[f"cam{k}"] = norm_err(g, want)
    print({k: f"{v:.1e}" for k, v in errs.items()})
    assert errs or filter == 0
    views_of = {f"tf{k}": [i for i, w in enumerate(VIEWS) if w[1] == k] for k in range(2)}
    views_of.update({f"vol{k}": [i for i, w in enumerate(VIEWS) if w[0] == k] for k in range(2)})
    views_of.update({f"cam{i}": [i] for i in range(len(VIEWS))})
    for k, e in errs.items():
        tol = 1e-4 if all(kinks[i] == 0 for i in views_of[k]) else 1e-3
        assert e <= tol, (k, e, tol)


@pytest.mark.parametrize("filter", [0, 1])
@pytest.mark.parametrize("tf_mode", [0, 1, 2, 3])
@pytest.mark.parametrize("blend", [0, 1])
def test_heterogeneous_batch_matches_oracle(filter, tf_mode, blend):
    seed = 7000 + 100 * filter + 10 * tf_mode + blend
    print(f"case filter={filter} tf_mode={tf_mode} blend={blend} seed={seed}", flush=True)
    case = Case(seed, filter, tf_mode, blend)
    plan = case.plan()
    plan.run()
    torch.cuda.synchronize()
    check_against_oracle(case)


@pytest.mark.parametrize("tf_mode", [0, 1, 2, 3])
@pytest.mark.parametrize("grads", [("tf", "volume"), ("tf", "camera"), ("camera",)])
def test_gradient_subsets_match_oracle(tf_mode, grads):
    """Plans with volume but no camera gradients run the volume instance (four waves per SIMD, chunk
    table sized beside the TF tables); plans with camera but no volume gradients run the TF + position
    instance.  Each is held to the oracle like the full-gradient instance."""
    seed = 7300 + 10 * tf_mode + len(grads)
    print(f"case tf_mode={tf_mode} grads={grads} seed={seed}", flush=True)
    case = Case(seed, 1, tf_mode, 0, grads=grads)
    plan = case.plan()
    print("chunk table slots", plan.chunk_slots)
    assert (plan.chunk_slots >= 32) if "volume" in grads else (plan.chunk_slots == 0)
    plan.run()
    torch.cuda.synchronize()
    check_against_oracle(case)


def test_large_tf_takes_the_smallest_chunk_table():
    """A 256-row texture TF and its gradient copies fill most of the LDS a volume-gradient workgroup
    may hold: the plan falls back to the smallest chunk table (32 slots, flushed most often).  The
    256-texel TF's slope amplifies FP32 density rounding 4x more than C2's 64 texels: the forward is
    held to 3x the FP32 CPU twin's error where that exceeds 1e-5."""
    case = Case(7400, 1, 1, 0, grads=("tf", "volume"), tf_rows=256)
    plan = case.plan(checkpoint_interval=8)
    assert plan.chunk_slots == 32, plan.chunk_slots
    plan.run()
    torch.cuda.synchronize()
    check_against_oracle(case, twin_bound=True)


def test_ill_conditioned_linear_tf_as_accurate_as_fp32_twin():
    """Control points 0.005 apart: FP32 density error is amplified ~1/gap by the TF slope.
    The GPU gradients must be as accurate as the FP32 CPU twin of the reference's
    mathematics (same sample counts, exact replay) is on the same inputs."""
    case = Case(7121, 1, 2, 1, spaced=False)
    assert min(np.diff(t[:, 4]).min() for t in case.tfs) < 0.01
    plan = case.plan()
    plan.run()
    torch.cuda.synchronize()
    terms = [v.termination.cpu().numpy() for v in case.views]
    _, g_vol, g_tf, g_cam = oracle_batch(case, terms, case.adj_colors)
    for i, (vi, ti, w, h, st) in enumerate(VIEWS):
        s = case.scene(i)
        t = terms[i][None]
        c32, _ = oracle.forward(s, dtype=np.float32, frozen_term=t)
        g32 = oracle.adjoint(s, c32, t, case.adj_colors[i][None], grads=("camera",), dtype=np.float32,
                             method=oracle.EXACT_REPLAY)
        twin = oracle.generate_rays_adjoint(case.vps[i][None].astype(np.float64), h, w, FOV,
                                            g32["ray_start"].astype(np.float64), g32["ray_dir"].astype(np.float64))[0]
        e_twin = norm_err(twin, g_cam[i])
        e_gpu = norm_err(case.g_cams[i].double().cpu().numpy(), g_cam[i])
        print(f"view {i}: camera gradient error gpu {e_gpu:.1e}, FP32 twin {e_twin:.1e}")
        assert e_gpu <= max(1e-4, 3.0 * e_twin), (i, e_gpu, e_twin)


def test_batched_forward_is_bitwise_the_tiled_kernel():
    """Same rays, same march: the batched kernel reproduces the single-batch tiled kernel exactly."""
    wl = scenes.workload_c2().to(dev())
    wl.width = wl.height = 72
    views = []
    for b in range(wl.viewports.shape[0]):
        views.append(bt.BatchView(volume=0, tf=0, width=wl.width, height=wl.height,
                                  camera=wl.viewports[b].contiguous(), fov_y=wl.fov_y, step=wl.step))
    plan = bt.BatchPlan([bt.BatchVolume(wl.volume[0].contiguous())], [bt.BatchTf(wl.tf[0].contiguous())], views,
                        volume_filter_mode=1, tf_mode=1, blend_mode=0, forward_only=True)
    plan.run()
    inp = pr.RendererInputs()
    inp.screen_size = pr.int2(wl.width, wl.height)
    inp.volume = wl.volume
    inp.volume_filter_mode = pr.VolumeFilterMode.Trilinear
    inp.box_min = pr.real3(-0.5, -0.5, -0.5)
    inp.box_size = pr.real3(1, 1, 1)
    inp.camera_mode = pr.CameraMode.ReferenceFrame
    inp.camera = pr.CameraReferenceFrame(wl.viewports, wl.fov_y)
    inp.step_size = wl.step
    inp.tf_mode = pr.TFMode.Texture
    inp.tf = wl.tf
    inp.blend_mode = pr.BlendMode.BeerLambert
    B = wl.viewports.shape[0]
    color = torch.empty(B, wl.height, wl.width, 4, device=dev())
    term = torch.empty(B, wl.height, wl.width, dtype=torch.int32, device=dev())
    pr.Renderer.render_forward(inp, pr.RendererOutputs(color, term), path="tiled")
    torch.cuda.synchronize()
    for b, v in enumerate(views):
        assert torch.equal(v.termination, term[b])
        assert torch.equal(v.color, color[b])


@pytest.mark.parametrize("loss", ["l1", "l2"])
def test_fused_loss_equals_given_adjoint(loss):
    """Loss gradient formed in registers == the same kernel fed torch's loss gradient."""
    case = Case(7400, 1, 1, 0, grads=("tf", "volume", "camera"), loss=loss)
    plan = case.plan()
    plan.run()
    torch.cuda.synchronize()
    colors = [v.color.clone() for v in case.views]
    targets = [v.target for v in case.views]
    N = sum(c.numel() for c in colors)
    cs = [c.clone().requires_grad_(True) for c in colors]
    if loss == "l1":
        ref = sum((c - t).abs().sum() for c, t in zip(cs, targets)) / N
    else:
        ref = sum(((c - t) ** 2).sum() for c, t in zip(cs, targets)) / N
    ref.backward()
    lv = float(case.loss_out.item())
    assert abs(lv - float(ref)) <= 1e-5 * abs(float(ref)), (lv, float(ref))
    g1 = ([g.clone() for g in case.g_tfs], [g.clone() for g in case.g_vols], [g.clone() for g in case.g_cams])
    # same batch, adjoint colour given
    case2 = Case(7400, 1, 1, 0, grads=("tf", "volume", "camera"), loss=None)
    for i, v in enumerate(case2.views):
        v.adj_color = cs[i].grad.contiguous()
    plan2 = case2.plan()
    plan2.run()
    torch.cuda.synchronize()
    for a, b in zip(g1[0] + g1[1] + g1[2], case2.g_tfs + case2.g_vols + case2.g_cams):
        e = norm_err(a.double().cpu().numpy(), b.double().cpu().numpy())
        assert e <= 1e-5, e


def test_autotune_keeps_results():
    case = Case(7500, 1, 3, 0)
    plan = case.plan()
    plan.run()
    torch.cuda.synchronize()
    ref_c = [v.color.clone() for v in case.views]
    ref_g = [g.clone() for g in case.g_tfs + case.g_vols + case.g_cams]
    chosen = plan.autotune(iterations=2)
    print("autotuned", chosen)
    assert chosen["checkpoint_interval"] in (4, 8, 16, 32, 64) and chosen["wave_width"] in (4, 8, 16, 32)
    assert all(float(g.abs().max()) == 0.0 for g in case.g_tfs + case.g_vols + case.g_cams)  # left zeroed
    plan.run()
    torch.cuda.synchronize()
    for a, v in zip(ref_c, case.views):
        assert torch.equal(a, v.color)  # per-ray arithmetic does not depend on tiles or K
    for a, b in zip(ref_g, case.g_tfs + case.g_vols + case.g_cams):
        assert norm_err(a.double().cpu().numpy(), b.double().cpu().numpy()) <= 1e-5


@pytest.mark.parametrize("grads", [("tf", "volume", "camera"), ("tf", "volume")])
def test_every_option_gives_the_same_gradients(grads):
    case = Case(7600, 1, 2, 1, grads=grads)
    plan = case.plan()
    plan.run()
    torch.cuda.synchronize()
    got = lambda: [g for g in case.g_tfs + case.g_vols + case.g_cams if g is not None]
    ref = [g.clone() for g in got()]
    for K in (4, 8, 16, 32, 64):
        for ww in (4, 8, 16, 32):
            for grid in (0, 8, 37):
                plan.set_options(K, ww, grid, density_store=(grid != 8))
                plan.run()
                torch.cuda.synchronize()
                for a, b in zip(ref, got()):
                    e = norm_err(a.double().cpu().numpy(), b.double().cpu().numpy())
                    assert e <= 1e-5, (K, ww, grid, e)


@pytest.mark.parametrize("filter,tf_mode,blend", [(1, 1, 0), (1, 3, 1), (0, 2, 0), (1, 0, 1)])
def test_density_store_matches_reinterpolation(filter, tf_mode, blend):
    """The reverse march reads the forward march's stored densities (default) or re-interpolates them:
    the same densities, so colours agree bitwise and gradients up to the order of atomic additions."""
    case = Case(7700 + 10 * filter + tf_mode, filter, tf_mode, blend)
    results = []
    for store in (True, False):
        plan = case.plan(checkpoint_interval=16, wave_width=8, grid=0, density_store=store)
        assert plan.options["density_store"] == store
        plan.run()
        torch.cuda.synchronize()
        results.append(([v.color.clone() for v in case.views],
                        [g.clone() for g in case.g_tfs + case.g_vols + case.g_cams if g is not None]))
    (c0, g0), (c1, g1) = results
    for a, b in zip(c0, c1):
        assert torch.equal(a, b)
    for a, b in zip(g0, g1):
        e = norm_err(a.double().cpu().numpy(), b.double().cpu().numpy())
        assert e <= 1e-6, e


def test_adam_matches_torch():
    torch.manual_seed(5)
    d = dev()
    shapes = [(40, 33), (7, 5), (1000,)]
    ours = [torch.rand(s, device=d) for s in shapes]
    refs = [p.clone().requires_grad_(True) for p in ours]
    grads = [torch.zeros_like(p) for p in ours]
    opt = bt.Adam([(ours[0], grads[0]), (ours[1], grads[1], 4, (0.0, 0.9)), (ours[2], grads[2])], lr=0.01)
    ref = torch.optim.Adam(refs, lr=0.01, foreach=False)
    for step in range(6):
        for g, r in zip(grads, refs):
            g.copy_(torch.randn_like(g) * (0.1 + step))
            r.grad = g.clone()
        # torch reference of the column-restricted, clamped block
        saved = refs[1].detach()[:, 4:].clone()
        ref.step()
        with torch.no_grad():
            refs[1][:, 4:] = saved
            refs[1][:, :4].clamp_(0.0, 0.9)
        opt.step()
        torch.cuda.synchronize()
    for k in range(3):
        e = (ours[k] - refs[k].detach()).abs().max().item()
        assert e <= 2e-6, (k, e)
    assert torch.equal(ours[1][:, 4:], refs[1].detach()[:, 4:])
    assert float(ours[1][:, :4].max()) <= 0.9 and float(ours[1][:, :4].min()) >= 0.0
    assert opt.step_count == 6


def test_graph_replays_match_eager_steps():
    """3 captured steps replayed 3 times == 9 eager steps (checked after every replay)."""
    n, reps = 3, 3

    def setup():
        case = Case(7700, 1, 2, 0, grads=("tf", "volume"), loss="l2")
        plan = case.plan()
        params = [(v, g, None, (0.0, 1.0)) for v, g in zip(case.t_vols, case.g_vols)]
        params += [(t, g, 4, None) for t, g in zip(case.t_tfs, case.g_tfs)]
        return case, plan, bt.Adam(params, lr=0.02)

    e_case, e_plan, e_opt = setup()
    eager = []
    for k in range(n * reps):
        e_plan.run()
        e_opt.step()
        if (k + 1) % n == 0:
            torch.cuda.synchronize()
            eager.append([x.clone() for x in e_case.t_vols + e_case.t_tfs])
    g_case, g_plan, g_opt = setup()
    g_plan.run()  # warm-up outside the capture
    g_plan.zero_gradients()
    torch.cuda.synchronize()
    steps = bt.GraphedSteps(g_plan, g_opt, n)
    for r in range(reps):
        steps.replay()
        torch.cuda.synchronize()
        # eager work between replays
        scratch = torch.randn(1 << 16, device=dev())
        _ = float(scratch.sum())
        for a, b in zip(eager[r], g_case.t_vols + g_case.t_tfs):
            e = (a - b).abs().max().item() / max(a.abs().max().item(), 1e-30)
            assert e <= 1e-4, (r, e)
    assert g_opt.step_count == n * reps
    loss_graph = float(g_case.loss_out.item())
    assert math.isfinite(loss_graph)
