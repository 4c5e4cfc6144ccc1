"""Pins the oracle's whole-image forward and adjoint (renderer_kernels.cuh:22-497).

* closed forms: empty rays, constant density under Beer-Lambert and alpha
  blending (colour, alpha, sample count, early termination, d/d(density));
* FP64 central differences of the oracle forward vs the oracle adjoint for
  volume, TF, per-pixel ray start/direction and step size, with the sample
  schedule frozen (tmin and the sample count of the unperturbed pass), which
  is exactly what the reference's adjoint differentiates (renderer_kernels.cuh:
  378-412 treats tmin/tmax and the count as constants);
* invariants: inversion adjoint == exact-replay adjoint, linearity in the
  output adjoint, broadcast gradient == sum over views, FP32 twin == FP64;
* camera ray generation and its adjoint (camera_impl.cu) against closed forms and FD.
"""
import math

import numpy as np
import pytest

from helpers import oracle, camera_scene, to_f32


def axis_scene(density, absorption, color, step, blend=0, W=4, H=3, size=(6, 5, 7)):
    """Rays along +x through a box [0,1]^3 starting at x=-1: t0 = 1, t1 = 2."""
    X, Y, Z = size
    ys = (np.arange(H) + 0.5) / H
    zs = (np.arange(W) + 0.5) / W
    start = np.zeros((1, H, W, 3))
    start[0, :, :, 0] = -1.0
    start[0, :, :, 1] = ys[:, None]
    start[0, :, :, 2] = zs[None, :]
    d = np.zeros((1, H, W, 3))
    d[..., 0] = 1.0
    return dict(W=W, H=H, filter=1, camera=0, tf_mode=0, blend=blend, early_out=1 - 1e-5, fov_y=0.0,
                volume=np.full((1, X, Y, Z), density), box_min=np.zeros((1, 3)), box_size=np.ones((1, 3)),
                step=np.full((1, 1, 1), step), tf=np.array([[[absorption, color]]]), ray_start=start, ray_dir=d)


def test_empty_rays():
    s = axis_scene(0.5, 2.0, 1.0, 0.1)
    s["ray_dir"][..., 0] = -1.0  # pointing away from the box
    c, t = oracle.forward(s)
    assert np.all(c == 0) and np.all(t == 0)
    g = oracle.adjoint(s, c, t, np.ones_like(c), grads=("tf", "volume", "camera", "step"))
    assert all(np.all(v == 0) for v in g.values())


@pytest.mark.parametrize("blend", [0, 1])
def test_constant_density_closed_form(blend):
    rho, a, cscale, step = 0.6, 1.7, 0.9, 0.07
    s = axis_scene(rho, a, cscale, step, blend=blend)
    c, t, t0 = oracle.forward(s, return_t0=True)
    np.testing.assert_allclose(t0, 1.0, rtol=1e-15)
    n = int(math.floor(1.0 / step + 1e-12)) + 1  # samples with t0 + i*step <= t1
    assert np.all(t == n)
    alpha = 1 - math.exp(-a * rho * step) if blend == 0 else min(1.0, a * rho * step)
    A = 1 - (1 - alpha) ** n
    np.testing.assert_allclose(c[..., 3], A, rtol=1e-12)
    np.testing.assert_allclose(c[..., 0], cscale * rho * A, rtol=1e-12)
    # d C_r / d rho summed over all voxels (the field stays constant)
    g = oracle.adjoint(s, c, t, np.tile([1.0, 0, 0, 0], c.shape[:-1] + (1,)), grads=("volume", "tf"))
    q = 1 - alpha
    dq = -a * step * q if blend == 0 else -a * step
    dA = -n * q ** (n - 1) * dq
    want = cscale * A + cscale * rho * dA
    np.testing.assert_allclose(g["volume"].sum() / (s["W"] * s["H"]), want, rtol=1e-10)
    # d C_r / d color-scale = rho * A (per pixel)
    np.testing.assert_allclose(g["tf"][0, 0, 1] / (s["W"] * s["H"]), rho * A, rtol=1e-10)


def test_early_termination_count():
    rho, a, step = 1.0, 40.0, 0.05
    s = axis_scene(rho, a, 1.0, step)
    c, t = oracle.forward(s)
    q = math.exp(-a * rho * step)
    # the test `A > 1-1e-5` happens before each sample: stop after the first n with 1-q^n > 1-1e-5
    n = next(k for k in range(1, 100) if 1 - q ** k > 1 - 1e-5)
    assert np.all(t == n)
    assert np.all(c[..., 3] > 1 - 1e-5)


def fd_check(scene, grads, rng, eps=1e-6, n_vol=25, tol=2e-6):
    """FP64 central differences of <adj, color> vs the oracle adjoint (frozen schedule).

    Parameters whose perturbation moves a sample across a kink of the
    piecewise definitions (texel / control-point boundaries, cell faces) give
    different quotients at eps and eps/4; those are skipped, and at least 80%
    of the probed parameters must be smooth.
    """
    c, t, t0 = oracle.forward(scene, return_t0=True)
    adj = rng.standard_normal(c.shape)
    g = oracle.adjoint(scene, c, t, adj, grads=grads, frozen_t0=t0)
    g_exact = oracle.adjoint(scene, c, t, adj, grads=grads, frozen_t0=t0, method=oracle.EXACT_REPLAY)

    def loss(sc):
        cc, _ = oracle.forward(sc, frozen_term=t, frozen_t0=t0)
        return float((cc * adj).sum())

    def fd(key, idx, e):
        sp = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in scene.items()}
        sm = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in scene.items()}
        sp[key][idx] += e
        sm[key][idx] -= e
        return (loss(sp) - loss(sm)) / (2 * e)

    probes = []
    if "volume" in grads:
        for f in rng.choice(scene["volume"].size, size=n_vol, replace=False):
            probes.append(("volume", "volume", np.unravel_index(f, scene["volume"].shape)))
    if "tf" in grads:
        for idx in np.ndindex(scene["tf"].shape):
            probes.append(("tf", "tf", idx))
    if "camera" in grads:
        for key in ("ray_start", "ray_dir"):
            for f in rng.choice(scene[key].size, size=12, replace=False):
                probes.append((key, key, np.unravel_index(f, scene[key].shape)))
    if "step" in grads:
        for f in rng.choice(scene["step"].size, size=min(12, scene["step"].size), replace=False):
            probes.append(("step", "step", np.unravel_index(f, scene["step"].shape)))
    rows = []
    for key, gk, idx in probes:
        f1, f2 = fd(key, idx, eps), fd(key, idx, eps / 4)
        rows.append((f1, f2, g[gk][idx], g_exact[gk][idx]))
    rows = np.array(rows)
    scale = np.abs(rows[:, 0]).max() + 1e-30
    smooth = np.abs(rows[:, 0] - rows[:, 1]) <= 1e-7 * scale
    assert smooth.mean() >= 0.8, f"only {smooth.sum()} of {len(rows)} probes are smooth"
    r = rows[smooth]
    assert np.abs(r[:, 2] - r[:, 0]).max() / scale < tol
    assert np.abs(r[:, 3] - r[:, 0]).max() / scale < tol
    return rows


@pytest.mark.parametrize("tf_mode", [0, 1, 2, 3])
@pytest.mark.parametrize("blend", [0, 1])
def test_render_fd_trilinear(tf_mode, blend):
    rng = np.random.default_rng(100 + 10 * tf_mode + blend)
    s = camera_scene(rng, 6, 5, tf_mode=tf_mode, blend=blend, camera=0, density_scale=0.8)
    if blend == 1:
        s["tf"][..., 0 if tf_mode == 0 else 3] *= 0.5  # keep absorption*step < 1, away from the alpha clamp
    s["step"] = np.full((1, s["H"], s["W"]), 0.03) + rng.uniform(-0.005, 0.005, (1, s["H"], s["W"]))
    fd_check(s, ("volume", "tf", "camera", "step"), rng)


@pytest.mark.parametrize("tf_mode", [1, 3])
def test_render_fd_nearest(tf_mode):
    rng = np.random.default_rng(7 + tf_mode)
    s = camera_scene(rng, 6, 5, filter=0, tf_mode=tf_mode, camera=0)
    fd_check(s, ("volume", "tf"), rng, eps=1e-7)


def test_inversion_equals_replay_near_opacity():
    rng = np.random.default_rng(11)
    s = camera_scene(rng, 8, 8, tf_mode=1, camera=2, B=2)
    s["tf"][..., 3] *= 8.0  # rays reach high opacity
    c, t = oracle.forward(s)
    assert c[..., 3].max() > 0.99
    adj = rng.standard_normal(c.shape)
    gi = oracle.adjoint(s, c, t, adj, grads=("tf", "volume"))
    ge = oracle.adjoint(s, c, t, adj, grads=("tf", "volume"), method=oracle.EXACT_REPLAY)
    for k in gi:
        assert np.linalg.norm(gi[k] - ge[k]) / np.linalg.norm(ge[k]) < 1e-7


def test_adjoint_linearity_and_broadcast_sum():
    rng = np.random.default_rng(12)
    s = camera_scene(rng, 7, 6, tf_mode=2, camera=2, B=3)
    c, t = oracle.forward(s)
    a1, a2 = rng.standard_normal(c.shape), rng.standard_normal(c.shape)
    g1 = oracle.adjoint(s, c, t, a1, grads=("tf", "volume"))
    g2 = oracle.adjoint(s, c, t, a2, grads=("tf", "volume"))
    g12 = oracle.adjoint(s, c, t, 2.0 * a1 - 0.5 * a2, grads=("tf", "volume"))
    for k in g1:
        np.testing.assert_allclose(g12[k], 2.0 * g1[k] - 0.5 * g2[k], rtol=1e-9, atol=1e-12)
    # volume/TF are broadcast (batch 1) over 3 views: gradient == sum of per-view gradients
    per = []
    for b in range(3):
        sb = dict(s)
        sb["cam_matrix"] = s["cam_matrix"][b:b + 1]
        per.append(oracle.adjoint(sb, c[b:b + 1], t[b:b + 1], a1[b:b + 1], grads=("tf", "volume")))
    for k in g1:
        np.testing.assert_allclose(g1[k], sum(p[k] for p in per), rtol=1e-9, atol=1e-12)
    # a per-view (batch B) gradient tensor receives each view separately
    gb = oracle.adjoint(s, c, t, a1, grads=("tf",), adj_tf_batch=3)
    for b in range(3):
        np.testing.assert_allclose(gb["tf"][b], per[b]["tf"][0], rtol=1e-9, atol=1e-12)


def test_fp32_twin_matches_fp64_forward():
    """The FP32 twin tracks FP64 to float rounding: 1e-5 on the smooth C1 workload;
    1e-4 on white-noise volumes, whose ~1/voxel gradients amplify the float
    rounding of sample positions."""
    from helpers import scene_from_workload, scenes
    s = scene_from_workload(scenes.workload_c1())
    c64, t64 = oracle.forward(s)
    c32, t32 = oracle.forward(s, dtype=np.float32)
    assert np.mean(t32 == t64) > 0.999
    c64f, _ = oracle.forward(s, frozen_term=t32)
    assert np.abs(c32 - c64f).max() / np.abs(c64f).max() < 1e-5
    rng = np.random.default_rng(13)
    for tf_mode in range(4):
        s = camera_scene(rng, 16, 12, tf_mode=tf_mode, camera=2, B=2)
        c64, t64 = oracle.forward(s)
        c32, t32 = oracle.forward(s, dtype=np.float32)
        assert np.mean(t32 == t64) > 0.99
        c64f, _ = oracle.forward(s, frozen_term=t32)
        assert np.abs(c32 - c64f).max() / max(np.abs(c64f).max(), 1e-30) < 1e-4


def test_generate_rays_closed_form_and_fd():
    rng = np.random.default_rng(14)
    from diffdvr_amd import scenes
    vp = scenes.sphere_viewports(rng.uniform(-1, 1, 2), rng.uniform(0, 6, 2), 2.5).double().numpy()
    H, W, fov = 5, 7, math.radians(45)
    rs, rd = oracle.generate_rays(vp, H, W, fov)
    np.testing.assert_allclose(rs, np.broadcast_to(vp[:, None, None, 0, :], rs.shape))
    np.testing.assert_allclose(np.linalg.norm(rd, axis=-1), 1.0, rtol=1e-14)
    # centre pixel looks along front = cross(up, right); the ray of an odd-sized image centre
    front = np.cross(vp[:, 2], vp[:, 1])
    np.testing.assert_allclose(rd[:, 2, 3], front / np.linalg.norm(front, axis=1, keepdims=True), atol=1e-14)
    # corner pixel spans tan(fov/2) (minus half a pixel) in y
    tanY = math.tan(fov / 2)
    fy = 2 * (0 + 0.5) / H - 1
    up_comp = np.einsum("bi,bi->b", rd[:, 0, 3], vp[:, 2])
    fr_comp = np.einsum("bi,bi->b", rd[:, 0, 3], front / np.linalg.norm(front, axis=1, keepdims=True))
    np.testing.assert_allclose(up_comp / fr_comp, fy * tanY, rtol=1e-7)  # frame is float32-rounded
    # adjoint vs FD
    a_s, a_d = rng.standard_normal(rs.shape), rng.standard_normal(rd.shape)
    g = oracle.generate_rays_adjoint(vp, H, W, fov, a_s, a_d)
    eps = 1e-6
    for idx in np.ndindex(vp.shape):
        p, m = vp.copy(), vp.copy()
        p[idx] += eps
        m[idx] -= eps
        lp = sum((x * y).sum() for x, y in zip(oracle.generate_rays(p, H, W, fov), (a_s, a_d)))
        lm = sum((x * y).sum() for x, y in zip(oracle.generate_rays(m, H, W, fov), (a_s, a_d)))
        assert g[idx] == pytest.approx((lp - lm) / (2 * eps), rel=1e-6, abs=1e-8)


def test_kink_census():
    """The census flags exactly the samples that sit on a kink of the reference's mathematics."""
    # constant density 0.6 along +x, 11 samples per ray (t = 1.0, 1.1, ..., 2.0)
    s = axis_scene(0.6, 1.7, 0.9, 0.1)
    _, t = oracle.forward(s)
    assert np.all(oracle.kink_census(s, t) == 0)        # identity TF, density inside (0, 1)
    # density on the identity clamp at 1: a uniform volume gives exactly 1 in every precision (no
    # kink); with y-slices 1, 1, 1, 0.9, 0.9 the middle row (p.y = 2) has density 1 from unequal
    # corners, the other rows lie in uniform cells (y = 0..1) or at 0.9
    s1 = axis_scene(1.0, 1.7, 0.9, 0.1)
    assert np.all(oracle.kink_census(s1, t) == 0)
    s1["volume"][:, :, 3:, :] = 0.9
    c1 = oracle.kink_census(s1, t)
    assert np.all(c1[0, 1] == t[0, 1]) and np.all(c1[0, [0, 2]] == 0)
    # linear TF with a control point exactly at the density 0.6: on the kink where the corners differ
    s2 = dict(s, tf_mode=2, tf=np.array([[[0.1, 0.2, 0.3, 1.0, 0.0], [0.4, 0.5, 0.6, 2.0, 0.6],
                                           [0.7, 0.8, 0.9, 3.0, 1.0]]]))
    assert np.all(oracle.kink_census(s2, t) == 0)
    s2["volume"] = s2["volume"].copy()
    s2["volume"][:, :, 3:, :] = 0.7
    c2 = oracle.kink_census(s2, t)
    assert np.all(c2[0, 1] == t[0, 1]) and np.all(c2[0, [0, 2]] == 0)
    s2["tf"][0, 1, 4] = 0.61                            # 0.01 away: no kink
    assert np.all(oracle.kink_census(s2, t) == 0)
    # texture TF with R = 5: texel edges where density * 5 - 0.5 is an integer, e.g. 0.5
    s3 = dict(axis_scene(0.5, 1.7, 0.9, 0.1), tf_mode=1, tf=np.ones((1, 5, 4)))
    assert np.all(oracle.kink_census(s3, t) == t)
    # alpha blending's clamp at absorption * step = 1: 0.6 * a * 0.1 = 1
    s4 = axis_scene(0.6, 1.0 / 0.06, 0.9, 0.1, blend=1)
    _, t4 = oracle.forward(s4, frozen_term=t)
    assert np.all(oracle.kink_census(s4, t) == t)
    # nearest filtering, X = 6 (voxel 0.2): world x = 0, 0.1, ..., 1 gives p.x = 0, 0.5, ..., 5, so
    # the samples at p.x = 0.5, 1.5, 2.5, 3.5, 4.5 lie on rounding edges; p.y = 0.67, 2, 3.33 and
    # p.z = 0.75, 2.25, 3.75, 5.25 never do
    s5 = dict(s, filter=0)
    assert np.all(oracle.kink_census(s5, t) == 5)
    assert np.all(oracle.kink_census(s5, t, eps_face=1e-6) == 5)  # faces are a trilinear kink only
    # trilinear cell faces on request: p.x = 1, ..., 5 on every ray (p.x = 0 is not a kink), and every
    # sample of the middle row, whose p.y = 2 lies on a face
    census = oracle.kink_census(s, t, eps_face=1e-6)
    assert np.all(census[0, [0, 2]] == 5) and np.all(census[0, 1] == t[0, 1])
    # pre-shaded volume: the alpha clamp reads the absorption channel, 10 * 0.1 = 1
    s6 = dict(s, filter=3, tf_mode=5, blend=1,
              volume=np.broadcast_to(np.array([0.6, 0.6, 0.6, 10.0])[:, None, None, None], (4, 6, 5, 7)).copy())
    assert np.all(oracle.kink_census(s6, t) == t)
    s6["volume"][3] = 9.0
    assert np.all(oracle.kink_census(s6, t) == 0)


def preshaded_scene(rng, W=6, H=5, blend=0):
    """A pre-shaded colour volume (r, g, b, absorption in the 4 batch slices) seen through per-pixel rays."""
    s = camera_scene(rng, W, H, filter=1, tf_mode=1, blend=blend, camera=0, smooth=True, X=12, Y=10, Z=14)
    vol = s["volume"][0]
    s["volume"] = to_f32(np.stack([vol * 0.9, 1.0 - vol, 0.5 + 0.4 * vol, vol * (4.0 if blend == 0 else 2.0)]))
    s["filter"], s["tf_mode"] = 3, 5
    return s


@pytest.mark.parametrize("blend", [0, 1])
def test_render_fd_preshaded(blend):
    """Pre-shaded volumes (renderer_interpolation.cuh:390-579): FP64 central differences vs both oracle
    adjoints for the 4-channel volume, per-pixel rays and step."""
    rng = np.random.default_rng(300 + blend)
    s = preshaded_scene(rng, blend=blend)
    s["step"] = np.full((1, s["H"], s["W"]), 0.03) + rng.uniform(-0.005, 0.005, (1, s["H"], s["W"]))
    assert oracle.batch_size(s) == 1
    fd_check(s, ("volume", "camera", "step"), rng)


def test_preshaded_constant_colour_closed_form():
    """A constant pre-shaded volume (r, g, b, a) under Beer-Lambert: colour = c * (1 - e^{-a dt n})."""
    s = preshaded_scene(np.random.default_rng(302))
    c = np.array([0.2, 0.5, 0.7, 3.0])
    s["volume"] = np.broadcast_to(c[:, None, None, None], s["volume"].shape).copy()
    col, term = oracle.forward(s)
    dt = float(s["step"].reshape(-1)[0])
    A = 1.0 - np.exp(-3.0 * dt * term)
    assert np.allclose(col[..., 3], A, atol=1e-12)
    assert np.allclose(col[..., :3], A[..., None] * c[None, None, None, :3], atol=1e-12)
