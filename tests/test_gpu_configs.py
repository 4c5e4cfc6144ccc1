"""The benchmark configurations at full size on the GPU, checked against the FP64 oracle.

Each test builds the configuration's whole persistent plan (every view of every
volume in one k_batched launch, as the benchmark runs it) and checks a subset
of its views against the oracle rendering that view alone, the way a reference
caller loops over Renderer.render_forward / render_adjoint per view
(pytests/diffdvr/renderer.py:127-213):
  * C3: 256^3 blobs + noise, identity TF x10, 3 Fibonacci views of 256^2, L2
    against a second volume, gradient w.r.t. every voxel;
  * C4: all 512 sphere cameras of sizes {64..512}^2 in one launch, Gaussian TF,
    L2 against perturbed-camera images; the latitude / longitude / radius
    gradients of one camera of every size against the oracle chain (its
    viewport_from_sphere, rays, ray adjoint, generateRays adjoint and the
    complex-step Jacobian of its viewport, camera.cpp:149-198,
    camera_impl.cu:302-357);
  * C5: the 16-volume plan (sizes 64..256^3, 4-32 views each); the density and
    TF gradients of the 256^3 volume 0 and the 64^3 volume 9 from two views
    each (every other view gets its own image as target, so it contributes an
    exact zero), then HIP Adam steps: the loss falls, and a captured graph of
    steps equals the same steps run eagerly.
Forward: termination index equal to the free-running oracle on >= 99.9% of
rays and never off by more than 1; colour (oracle replaying the GPU's sample
counts) <= 1e-5 of the image maximum.  Gradients normwise <= 1e-4; the kink
census (samples within FP32 rounding of a kink of the reference's derivative,
where FP32 may take the other side) must flag at most 0.1% of the samples.
"""
import math

import numpy as np
import pytest
import torch

from helpers import oracle, scenes
from gpu_helpers import dev, norm_err
from diffdvr_amd import batch as bt
from diffdvr_amd import pyrenderer as pr

pytestmark = pytest.mark.gpu


def view_scene(vol, tf, tf_mode, W, H, rays, step, box_min=(-0.5, -0.5, -0.5), box_size=(1, 1, 1)):
    """Oracle scene of one view with per-pixel rays (float64 arrays)."""
    rs, rd = rays
    return dict(W=W, H=H, filter=1, camera=0, tf_mode=tf_mode, blend=0, early_out=1 - 1e-5, fov_y=scenes.FOV_Y,
                volume=vol[None], box_min=np.array([box_min], dtype=np.float64),
                box_size=np.array([box_size], dtype=np.float64), step=np.full((1, 1, 1), step), tf=tf[None],
                ray_start=rs, ray_dir=rd)


def check_forward(s, color, term, what):
    """Termination vs the free-running oracle, colour vs the oracle replaying the GPU's counts."""
    t = term.cpu().numpy()[None]
    _, t_free = oracle.forward(s)
    match = float((t_free == t).mean())
    assert match >= 0.999 and np.abs(t_free - t).max() <= 1, (what, match)
    c_ora, _ = oracle.forward(s, frozen_term=t)
    census = oracle.kink_census(s, t)[0]
    kinked = int((census > 0).sum())
    assert census.sum() <= 1e-3 * t.sum(), (what, "kink census", int(census.sum()), int(t.sum()))
    c = color.double().cpu().numpy()[None]
    # trilinear densities: the colour is continuous at every kink of the derivative, so all rays count
    err = np.abs(c - c_ora).max() / max(np.abs(c_ora).max(), 1e-30)
    print(f"{what}: term match {match:.5f}, colour err {err:.2e}, kink samples {int(census.sum())} "
          f"of {int(t.sum())} on {kinked} rays", flush=True)
    assert err <= 1e-5, (what, err)
    return t, c_ora, kinked


# ---------------------------------------------------------------------------
# C3
# ---------------------------------------------------------------------------
def test_c3_volume_gradient_256():
    print("C3: 256^3, 3 views 256^2, volume gradient through k_batched", flush=True)
    d = dev()
    wl = scenes.workload_c3(device=d)
    pick = [0, 21, 45]
    vol = wl.volume[0].contiguous()
    tf = wl.tf[0].contiguous()
    W = H = wl.width
    vps = [wl.viewports[i].contiguous() for i in pick]
    tviews = [bt.BatchView(volume=0, tf=0, width=W, height=H, camera=v, fov_y=wl.fov_y, step=wl.step) for v in vps]
    bt.BatchPlan([bt.BatchVolume(wl.target_volume[0].contiguous())], [bt.BatchTf(tf)], tviews, volume_filter_mode=1,
                 tf_mode=0, forward_only=True).run()
    targets = [v.color.clone() for v in tviews]
    g_vol = torch.zeros_like(vol)
    views = [bt.BatchView(volume=0, tf=0, width=W, height=H, camera=v, fov_y=wl.fov_y, step=wl.step, target=t)
             for v, t in zip(vps, targets)]
    plan = bt.BatchPlan([bt.BatchVolume(vol, grad=g_vol)], [bt.BatchTf(tf)], views, volume_filter_mode=1, tf_mode=0,
                        blend_mode=0, loss="l2")
    plan.run()
    torch.cuda.synchronize()
    scale = 1.0 / (len(pick) * W * H * 4)
    vol64 = vol.double().cpu().numpy()
    tf64 = tf.double().cpu().numpy()
    g_all = g_vol.double().cpu().numpy()
    scenes_, want_all, want_clean, twin_all = [], np.zeros_like(vol64), np.zeros_like(vol64), np.zeros_like(vol64)
    for k, (v, view) in enumerate(zip(vps, views)):
        rays = oracle.generate_rays(v.double().cpu().numpy()[None], H, W, wl.fov_y)
        s = view_scene(vol64, tf64, 0, W, H, rays, wl.step)
        t, c_ora, _ = check_forward(s, view.color, view.termination, f"C3 view {pick[k]}")
        c = view.color.double().cpu().numpy()[None]
        adj = 2.0 * (c - targets[k].double().cpu().numpy()[None]) * scale
        want_all += oracle.adjoint(s, c_ora, t, adj, grads=("volume",), method=oracle.EXACT_REPLAY)["volume"][0]
        # the FP32 CPU twin of the reference (same sample counts) on all rays
        c32, _ = oracle.forward(s, dtype=np.float32, frozen_term=t)
        twin_all += oracle.adjoint(s, c32, t, adj, grads=("volume",), dtype=np.float32,
                                   method=oracle.EXACT_REPLAY)["volume"][0]
        # rays with a sample within FP32 rounding of a kink of the identity TF's clamp: their loss
        # gradient is zeroed on both sides (target := the rendered colour)
        kinked = oracle.kink_census(s, t)[0] > 0
        targets[k][torch.as_tensor(kinked, device=d)] = view.color[torch.as_tensor(kinked, device=d)]
        adj[0][kinked] = 0.0
        want_clean += oracle.adjoint(s, c_ora, t, adj, grads=("volume",), method=oracle.EXACT_REPLAY)["volume"][0]
    plan.run()
    torch.cuda.synchronize()
    got = g_vol.double().cpu().numpy()
    e_clean = norm_err(got[2:-2, 2:-2, 2:-2], want_clean[2:-2, 2:-2, 2:-2])
    e_all = norm_err(g_all[2:-2, 2:-2, 2:-2], want_all[2:-2, 2:-2, 2:-2])
    e_twin = norm_err(twin_all[2:-2, 2:-2, 2:-2], want_all[2:-2, 2:-2, 2:-2])
    print(f"C3 volume gradient (interior): kink-free rays {e_clean:.2e}; all rays {e_all:.2e} "
          f"(FP32 CPU twin of the reference on all rays: {e_twin:.2e})", flush=True)
    assert e_clean <= 1e-4, e_clean
    assert e_all <= max(1e-4, 3.0 * e_twin), (e_all, e_twin)


# ---------------------------------------------------------------------------
# C4
# ---------------------------------------------------------------------------
def test_c4_camera_gradients_512_cameras():
    print("C4: 512 cameras of sizes {64..512}^2 in one launch, camera gradients", flush=True)
    d = dev()
    wl = scenes.workload_c4()
    cams, tcams = wl.cameras[0], wl.target_cameras[0]
    N = len(cams.sizes)
    vol = wl.volumes[0].to(d).contiguous()
    tf = wl.tfs[0].to(d).contiguous()
    step = wl.steps[0]
    tvp = tcams.viewports(d)
    tviews = [bt.BatchView(volume=0, tf=0, width=w, height=h, camera=tvp[i], fov_y=wl.fov_y, step=step)
              for i, (w, h) in enumerate(cams.sizes)]
    bt.BatchPlan([bt.BatchVolume(vol)], [bt.BatchTf(tf)], tviews, volume_filter_mode=1, tf_mode=3,
                 forward_only=True).run()
    targets = [v.color for v in tviews]
    lat = cams.lat.to(d).float().requires_grad_(True)
    lon = cams.lon.to(d).float().requires_grad_(True)
    rad = cams.radius.to(d).float().requires_grad_(True)
    vp = pr.Camera.viewport_from_sphere(torch.zeros(N, 3, device=d), lon, lat, rad, pr.Orientation.Yp)
    vp_buf = vp.detach().contiguous()
    g_vp = torch.zeros(N, 3, 3, device=d)
    views = [bt.BatchView(volume=0, tf=0, width=w, height=h, camera=vp_buf[i], fov_y=wl.fov_y, step=step,
                          target=targets[i], camera_grad=g_vp[i]) for i, (w, h) in enumerate(cams.sizes)]
    plan = bt.BatchPlan([bt.BatchVolume(vol)], [bt.BatchTf(tf)], views, volume_filter_mode=1, tf_mode=3,
                        blend_mode=0, loss="l2")
    plan.run()
    vp.backward(g_vp)
    torch.cuda.synchronize()
    n_all = float(sum(w * h for w, h in cams.sizes) * 4)
    sizes = np.array([w for w, _ in cams.sizes])
    pick = [int(np.nonzero(sizes == s)[0][0]) for s in (64, 128, 256, 512)]
    vol64 = vol.double().cpu().numpy()
    tf64 = tf.double().cpu().numpy()
    lat32, lon32, rad32 = (x.detach().double().cpu().numpy() for x in (lat, lon, rad))
    for i in pick:
        w, h = cams.sizes[i]
        # oracle chain from the float32 parameters the GPU used (yaw = longitude, pitch = latitude)
        args = (np.zeros((1, 3)), lon32[i:i + 1], lat32[i:i + 1], rad32[i:i + 1], pr.Orientation.Yp)
        vp64 = oracle.viewport_from_sphere(*args)
        vp32 = vp_buf[i].double().cpu().numpy()[None]
        assert np.abs(vp32 - vp64).max() <= 2e-6, np.abs(vp32 - vp64).max()   # the product's float32 viewport
        # rays of the viewport the GPU rendered (float32), so that colours compare at 1e-5
        rays = oracle.generate_rays(vp32, h, w, wl.fov_y)
        s = view_scene(vol64, tf64, 3, w, h, rays, step)
        t, c_ora, _ = check_forward(s, views[i].color, views[i].termination, f"C4 camera {i} ({w}^2)")
        adj = 2.0 * (views[i].color.double().cpu().numpy()[None] - targets[i].double().cpu().numpy()[None]) / n_all
        g = oracle.adjoint(s, c_ora, t, adj, grads=("camera",), method=oracle.EXACT_REPLAY)
        avp = oracle.generate_rays_adjoint(vp32, h, w, wl.fov_y, g["ray_start"], g["ray_dir"])
        J = oracle.viewport_from_sphere_jacobian(*args)[0]              # 3 x 3 x (yaw, pitch, distance)
        want_vp = avp[0]
        e_vp = norm_err(g_vp[i].double().cpu().numpy(), want_vp)
        want = np.sum(J * want_vp[..., None], axis=(0, 1))             # d/d lon, lat, r
        got = np.array([lon.grad[i, 0].item(), lat.grad[i, 0].item(), rad.grad[i, 0].item()])
        e = norm_err(got, want)
        print(f"C4 camera {i} ({w}^2): viewport grad err {e_vp:.2e}, lon/lat/r grad err {e:.2e}", flush=True)
        assert e_vp <= 1e-4 and e <= 1e-4, (i, e_vp, e)


# ---------------------------------------------------------------------------
# C5
# ---------------------------------------------------------------------------
def c5_plan(d, checked=None):
    """The C5 plan on the device: volumes, targets, TFs, views, gradients, loss.  With `checked`
    ({volume: [view positions]}), every other view gets its own rendering as target."""
    wl = scenes.workload_c5(device=d)
    vols = [v.contiguous() for v in wl.volumes]
    tgts = [v.contiguous() for v in wl.target_volumes]
    tfs = [t.to(d).contiguous() for t in wl.tfs]
    vps, sizes, vidx, steps, pos = [], [], [], [], []
    for k, cams in enumerate(wl.cameras):
        vp = cams.viewports(d)
        for i in range(len(cams.sizes)):
            vps.append(vp[i].contiguous())
            sizes.append(cams.sizes[i])
            vidx.append(k)
            steps.append(wl.steps[k])
            pos.append(i)

    def views_for(targets=None):
        return [bt.BatchView(volume=vidx[i], tf=vidx[i], width=sizes[i][0], height=sizes[i][1], camera=vps[i],
                             fov_y=wl.fov_y, step=steps[i], target=None if targets is None else targets[i])
                for i in range(len(vps))]

    tv = views_for()
    bt.BatchPlan([bt.BatchVolume(v) for v in tgts], [bt.BatchTf(t) for t in tfs], tv, volume_filter_mode=1, tf_mode=2,
                 forward_only=True).run()
    targets = [v.color for v in tv]
    if checked is not None:
        own = views_for()
        bt.BatchPlan([bt.BatchVolume(v) for v in vols], [bt.BatchTf(t) for t in tfs], own, volume_filter_mode=1,
                     tf_mode=2, forward_only=True).run()
        targets = [targets[i] if pos[i] in checked.get(vidx[i], ()) else own[i].color for i in range(len(vps))]
    g_vols = [torch.zeros_like(v) for v in vols]
    g_tfs = [torch.zeros_like(t) for t in tfs]
    views = views_for(targets)
    loss = torch.zeros(1, device=d)
    plan = bt.BatchPlan([bt.BatchVolume(v, grad=g) for v, g in zip(vols, g_vols)],
                        [bt.BatchTf(t, g) for t, g in zip(tfs, g_tfs)], views, volume_filter_mode=1, tf_mode=2,
                        blend_mode=0, loss="l2", loss_out=loss)
    return dict(wl=wl, vols=vols, tfs=tfs, views=views, targets=targets, vidx=vidx, pos=pos, vps=vps, sizes=sizes,
                steps=steps, g_vols=g_vols, g_tfs=g_tfs, loss=loss, plan=plan)


def test_c5_two_volumes_against_oracle():
    print("C5: 16-volume plan, density + TF gradients of volumes 0 (256^3) and 9 (64^3)", flush=True)
    d = dev()
    checked = {0: (0, 1), 9: (0, 1)}
    c = c5_plan(d, checked)
    assert tuple(c["vols"][0].shape) == (256, 256, 256) and tuple(c["vols"][9].shape) == (64, 64, 64)
    c["plan"].run()
    torch.cuda.synchronize()
    n_all = float(sum(w * h for w, h in c["sizes"]) * 4)
    got_all = {vi: (c["g_vols"][vi].double().cpu().numpy(), c["g_tfs"][vi].double().cpu().numpy()) for vi in checked}
    want = {}
    for vi, ps in checked.items():
        vol64 = c["vols"][vi].double().cpu().numpy()
        tf64 = c["tfs"][vi].double().cpu().numpy()
        acc = {k: np.zeros_like(vol64 if k[0] == "v" else tf64) for k in ("v_all", "t_all", "v_clean", "t_clean",
                                                                           "v_twin")}
        for i in range(len(c["views"])):
            if c["vidx"][i] != vi or c["pos"][i] not in ps:
                continue
            w, h = c["sizes"][i]
            rays = oracle.generate_rays(c["vps"][i].double().cpu().numpy()[None], h, w, c["wl"].fov_y)
            s = view_scene(vol64, tf64, 2, w, h, rays, c["steps"][i])
            view = c["views"][i]
            t, c_ora, _ = check_forward(s, view.color, view.termination, f"C5 volume {vi} view {c['pos'][i]} ({w}^2)")
            col = view.color.double().cpu().numpy()[None]
            adj = 2.0 * (col - c["targets"][i].double().cpu().numpy()[None]) / n_all
            g = oracle.adjoint(s, c_ora, t, adj, grads=("volume", "tf"), method=oracle.EXACT_REPLAY)
            acc["v_all"] += g["volume"][0]
            acc["t_all"] += g["tf"][0]
            c32, _ = oracle.forward(s, dtype=np.float32, frozen_term=t)
            acc["v_twin"] += oracle.adjoint(s, c32, t, adj, grads=("volume",), dtype=np.float32,
                                            method=oracle.EXACT_REPLAY)["volume"][0]
            # rays with a sample within FP32 rounding of a control point: loss gradient zeroed on
            # both sides (target := the rendered colour)
            kinked = oracle.kink_census(s, t)[0] > 0
            kt = torch.as_tensor(kinked, device=d)
            c["targets"][i][kt] = view.color[kt]
            adj[0][kinked] = 0.0
            g = oracle.adjoint(s, c_ora, t, adj, grads=("volume", "tf"), method=oracle.EXACT_REPLAY)
            acc["v_clean"] += g["volume"][0]
            acc["t_clean"] += g["tf"][0]
        want[vi] = acc
    c["plan"].run()
    torch.cuda.synchronize()
    inner = (slice(2, -2),) * 3
    for vi in checked:
        a = want[vi]
        e_vol = norm_err(c["g_vols"][vi].double().cpu().numpy()[inner], a["v_clean"][inner])
        e_tf = norm_err(c["g_tfs"][vi].double().cpu().numpy(), a["t_clean"])
        e_vol_all = norm_err(got_all[vi][0][inner], a["v_all"][inner])
        e_tf_all = norm_err(got_all[vi][1], a["t_all"])
        e_twin = norm_err(a["v_twin"][inner], a["v_all"][inner])
        print(f"C5 volume {vi}: kink-free rays: density grad err {e_vol:.2e}, TF grad err {e_tf:.2e}; all rays: "
              f"density {e_vol_all:.2e} (FP32 CPU twin {e_twin:.2e}), TF {e_tf_all:.2e}", flush=True)
        assert e_vol <= 1e-4 and e_tf <= 1e-4, (vi, e_vol, e_tf)
        assert e_vol_all <= max(1e-4, 3.0 * e_twin) and e_tf_all <= 1e-4, (vi, e_vol_all, e_twin, e_tf_all)
    # every other volume's views reproduce their targets exactly: zero gradient
    for vi in range(len(c["vols"])):
        if vi not in checked:
            assert float(c["g_vols"][vi].abs().max()) == 0.0 and float(c["g_tfs"][vi].abs().max()) == 0.0, vi


def test_c5_adam_loss_falls_and_graph_equals_eager():
    print("C5: HIP Adam steps on the 16-volume plan, eager and captured", flush=True)
    d = dev()

    def setup():
        c = c5_plan(d)
        params = [(v, g, None, (0.0, 1.0)) for v, g in zip(c["vols"], c["g_vols"])]
        params += [(t, g, 4, (0.0, 1e3)) for t, g in zip(c["tfs"], c["g_tfs"])]
        return c, bt.Adam(params, lr=1e-2)

    c, adam = setup()
    losses = []
    for _ in range(4):
        c["plan"].run()
        adam.step()
        losses.append(float(c["loss"].item()))
    print("eager losses", losses, flush=True)
    assert all(math.isfinite(x) for x in losses)
    assert losses[-1] < losses[0] and all(b <= a * 1.001 for a, b in zip(losses, losses[1:])), losses
    eager = [x.clone() for x in c["vols"] + c["tfs"]]
    del c, adam
    torch.cuda.empty_cache()
    g, gadam = setup()
    g["plan"].run()          # warm-up outside the capture
    g["plan"].zero_gradients()
    torch.cuda.synchronize()
    steps = bt.GraphedSteps(g["plan"], gadam, 2)
    for _ in range(2):
        steps.replay()
        torch.cuda.synchronize()
        _ = float(torch.randn(1 << 16, device=d).sum())   # eager work between replays
    assert gadam.step_count == 4
    # the loss of the last captured step's forward pass is the eager one's
    lg = float(g["loss"].item())
    assert abs(lg - losses[-1]) <= 1e-5 * abs(losses[-1]), (lg, losses[-1])
    # Gradients are sums of float atomics in arbitrary order, so they differ from run to run in
    # the last bits; Adam's m / sqrt(v) turns that into a visible difference only where a voxel's
    # contributions cancel.  Every parameter tensor must agree to 1e-4 on all but 1e-5 of its entries.
    for a, b in zip(eager, g["vols"] + g["tfs"]):
        bad = ((a - b).abs() > 1e-4 * max(a.abs().max().item(), 1e-30)).float().mean().item()
        assert bad <= 1e-5, bad


# ---------------------------------------------------------------------------
# central differences at configuration scale
# ---------------------------------------------------------------------------
def _slab(start, d, lo, hi):
    """Ray parameter interval of per-ray start/dir (N x 3) inside the box [lo, hi]."""
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = 1.0 / d
        a, b = (lo - start) * inv, (hi - start) * inv
    return np.max(np.minimum(a, b), axis=1), np.min(np.maximum(a, b), axis=1)


def test_c3_central_differences_256_voxels():
    """C3 (256^3 blobs + noise, identity TF x10, view 0 at 256^2, L2): the GPU's volume gradient
    at 256 random voxels against FP64 central differences of the oracle's loss with the
    reference's frozen sample schedule (renderer_kernels.cuh:378-412).  A voxel changes only the
    rays that pass through the 2x2x2 cells around it, so each quotient renders just those rays."""
    print("C3 central differences: 256 voxels of a 256^3 volume, one 256^2 view", flush=True)
    d = dev()
    wl = scenes.workload_c3(device=d, views=64)
    N = 256
    vol = wl.volume[0].contiguous()
    tf = wl.tf[0].contiguous()
    vp = wl.viewports[0].contiguous()
    W = H = wl.width
    tv = [bt.BatchView(volume=0, tf=0, width=W, height=H, camera=vp, fov_y=wl.fov_y, step=wl.step)]
    bt.BatchPlan([bt.BatchVolume(wl.target_volume[0].contiguous())], [bt.BatchTf(tf)], tv, volume_filter_mode=1,
                 tf_mode=0, forward_only=True).run()
    target = tv[0].color.clone()
    g_vol = torch.zeros_like(vol)
    view = bt.BatchView(volume=0, tf=0, width=W, height=H, camera=vp, fov_y=wl.fov_y, step=wl.step, target=target)
    bt.BatchPlan([bt.BatchVolume(vol, grad=g_vol)], [bt.BatchTf(tf)], [view], volume_filter_mode=1, tf_mode=0,
                 loss="l2").run()
    torch.cuda.synchronize()
    scale = 1.0 / (W * H * 4)
    vol64 = vol.double().cpu().numpy()
    g = g_vol.double().cpu().numpy()
    tg = target.double().cpu().numpy().reshape(-1, 4)
    term = view.termination.cpu().numpy().reshape(-1)
    rs, rd = oracle.generate_rays(vp.double().cpu().numpy()[None], H, W, wl.fov_y)
    rs, rd = rs.reshape(-1, 3), rd.reshape(-1, 3)
    s_full = view_scene(vol64, tf.double().cpu().numpy(), 0, W, H, (rs.reshape(1, H, W, 3), rd.reshape(1, H, W, 3)),
                        wl.step)
    _, _, t0 = oracle.forward(s_full, frozen_term=term.reshape(1, H, W), return_t0=True)
    t0 = t0.reshape(-1)
    t_last = t0 + (term - 1) * wl.step
    # voxels with a clear gradient whose 3x3x3 neighbourhood stays inside the identity TF's (0, 1)
    rng = np.random.default_rng(2024)
    inner = np.zeros_like(vol64, dtype=bool)
    inner[2:-2, 2:-2, 2:-2] = True
    lo3 = np.minimum.reduce([np.roll(vol64, sh, axis=ax) for ax in range(3) for sh in (-1, 0, 1)])
    hi3 = np.maximum.reduce([np.roll(vol64, sh, axis=ax) for ax in range(3) for sh in (-1, 0, 1)])
    ok = inner & (lo3 > 0.05) & (hi3 < 0.95) & (np.abs(g) > 1e-2 * np.abs(g).max())
    cand = np.argwhere(ok)
    assert len(cand) >= 256, len(cand)
    pick = cand[rng.choice(len(cand), 256, replace=False)]
    h = 1e-3
    voxel = 1.0 / (N - 1)
    fd, got = [], []
    for v in pick:
        lo = -0.5 + (v - 1) * voxel - 1e-9
        hi = -0.5 + (v + 1) * voxel + 1e-9
        a, b = _slab(rs, rd, lo, hi)
        hit = np.nonzero((term > 0) & (a <= b) & (b >= t0) & (a <= t_last))[0]
        s = view_scene(vol64, s_full["tf"][0], 0, len(hit), 1, (rs[hit][None, None], rd[hit][None, None]), wl.step)
        ft = term[hit].reshape(1, 1, -1)

        def loss():
            c, _ = oracle.forward(s, frozen_term=ft)
            return float(np.sum((c.reshape(-1, 4) - tg[hit]) ** 2) * scale)

        x = tuple(v)
        v0 = vol64[x]
        vol64[x] = v0 + h
        lp = loss()
        vol64[x] = v0 - h
        lm = loss()
        vol64[x] = v0
        fd.append((lp - lm) / (2 * h))
        got.append(g[x])
    fd, got = np.array(fd), np.array(got)
    e = norm_err(got, fd)
    print(f"C3 FD: GPU volume gradient vs FP64 central differences at 256 voxels: {e:.2e}", flush=True)
    assert e <= 1e-4, e


def test_c2_central_differences_tf():
    """C2 (128^3 blobs, 64x4 texture TF, L1 vs a second TF's images, view 0 at 256^2): the GPU's
    TF gradient of the loss of a 64x64 pixel window (the adjoint colour is zero elsewhere) against
    FP64 central differences of the oracle's loss over that window, for all 256 TF entries.
    Entries whose quotients at h and h/4 disagree (a sample's texel changes within h) are left
    out; at least 80% must remain."""
    print("C2 central differences: all 256 TF entries, 64x64 window of a 256^2 view", flush=True)
    d = dev()
    wl = scenes.workload_c2(device=d, views=8, size=256)
    W = H = 256
    x0, y0, n = 96, 96, 64
    vp = wl.viewports[0].contiguous()
    vol = wl.volume[0].contiguous()
    tv = [bt.BatchView(volume=0, tf=0, width=W, height=H, camera=vp, fov_y=wl.fov_y, step=wl.step)]
    bt.BatchPlan([bt.BatchVolume(vol)], [bt.BatchTf(wl.target_tf[0].contiguous())], tv, volume_filter_mode=1,
                 tf_mode=1, forward_only=True).run()
    target = tv[0].color.clone()
    fv = [bt.BatchView(volume=0, tf=0, width=W, height=H, camera=vp, fov_y=wl.fov_y, step=wl.step)]
    tf = wl.tf[0].contiguous()
    bt.BatchPlan([bt.BatchVolume(vol)], [bt.BatchTf(tf)], fv, volume_filter_mode=1, tf_mode=1, forward_only=True).run()
    torch.cuda.synchronize()
    scale = 1.0 / (n * n * 4)
    adj = torch.zeros(H, W, 4, device=d)
    win = (slice(y0, y0 + n), slice(x0, x0 + n))
    adj[win] = torch.sign(fv[0].color[win] - target[win]) * scale
    g_tf = torch.zeros_like(tf)
    view = bt.BatchView(volume=0, tf=0, width=W, height=H, camera=vp, fov_y=wl.fov_y, step=wl.step, adj_color=adj)
    bt.BatchPlan([bt.BatchVolume(vol)], [bt.BatchTf(tf, g_tf)], [view], volume_filter_mode=1, tf_mode=1).run()
    torch.cuda.synchronize()
    rs, rd = oracle.generate_rays(vp.double().cpu().numpy()[None], H, W, wl.fov_y)
    rs, rd = rs[:, y0:y0 + n, x0:x0 + n].copy(), rd[:, y0:y0 + n, x0:x0 + n].copy()
    tf64 = tf.double().cpu().numpy()
    s = view_scene(vol.double().cpu().numpy(), tf64, 1, n, n, (rs, rd), wl.step)
    term = view.termination[win].cpu().numpy()[None]
    sign = np.sign(view.color[win].double().cpu().numpy() - target[win].double().cpu().numpy())
    tg = target[win].double().cpu().numpy()

    def loss():
        c, _ = oracle.forward(s, frozen_term=term)
        return float(np.sum(np.abs(c[0] - tg)) * scale)

    rows = []
    for idx in np.ndindex(tf64.shape):
        q = []
        for hh in (1e-4, 2.5e-5):
            v0 = tf64[idx]
            tf64[idx] = v0 + hh
            lp = loss()
            tf64[idx] = v0 - hh
            lm = loss()
            tf64[idx] = v0
            q.append((lp - lm) / (2 * hh))
        rows.append((q[0], q[1], float(g_tf[idx].item())))
    rows = np.array(rows)
    sc = np.abs(rows[:, 0]).max()
    smooth = np.abs(rows[:, 0] - rows[:, 1]) <= 1e-6 * sc
    assert smooth.mean() >= 0.8, smooth.mean()
    e = norm_err(rows[smooth, 2], rows[smooth, 0])
    # the kernel differentiates |c - tg| with the sign of its own FP32 colour: count the pixels where
    # the oracle's FP64 colour has the other sign (each one shifts the gradient slightly)
    c64, _ = oracle.forward(s, frozen_term=term)
    flips = int(np.sum(np.sign(c64[0] - tg) != sign))
    print(f"L1 sign flips between FP32 and FP64 colours: {flips}", flush=True)
    print(f"C2 FD: GPU TF gradient vs FP64 central differences on {int(smooth.sum())} of 256 entries: {e:.2e}",
          flush=True)
    assert e <= 1e-4, e
