"""Auxiliary kernels on the GPU against the FP64 oracle: TF pre-shading and its
density adjoint for every TF mode incl. GaussianLog (renderer/tf_utils_impl.cu:17-190),
best-fit inversion (:225-520), importance sampling (renderer/losses_impl.cu:254-421)
and the elementwise log losses (:35-240), through pyrenderer.

Bounds: pre-shaded colours within 2e-6 of max|colour| (GaussianLog: 4e-6 relative
to max(1, |log value|)), density adjoints within 1e-5 of max|adj|;
best fit: the GPU density is a candidate of that voxel's stream whose FP64 cost is
within 1e-5 (relative) of the oracle's minimum; importance sampling: maxima within
1e-5 relative, decisions equal wherever the oracle's margin exceeds 1e-5; losses
within 1e-6 relative (float32) and 1e-12 (float64)."""
import numpy as np
import pytest
import torch

from helpers import oracle
from gpu_helpers import dev, t32
from diffdvr_amd import pyrenderer as pr
from test_oracle_aux import gaussian_tf, linear_tf

pytestmark = pytest.mark.gpu

TF_ENUM = {0: pr.TFMode.Identity, 1: pr.TFMode.Texture, 2: pr.TFMode.Linear, 3: pr.TFMode.Gaussian,
           6: pr.TFMode.GaussianLog}


def tf_for(mode, rng):
    if mode == 0:
        return np.array([[2.5, 0.7]])
    if mode == 1:
        return rng.uniform(0.0, 1.0, (33, 4))
    if mode == 2:
        return linear_tf(rng, R=7)
    return gaussian_tf(rng, R=5)


def x_fastest(a: np.ndarray) -> torch.Tensor:
    """A (C, X, Y, Z) array on the device with x the fastest axis (pyrenderer.Volume's layout)."""
    t = torch.as_tensor(np.ascontiguousarray(np.transpose(a, (0, 3, 2, 1))), dtype=torch.float32, device=dev())
    return t.permute(0, 3, 2, 1)


@pytest.mark.parametrize("shape", [(37, 29, 23), (32, 21, 24)])  # innermost extent % 4: scalar / vector lanes
@pytest.mark.parametrize("layout", ["contiguous", "x_fastest"])
@pytest.mark.parametrize("mode", [0, 1, 2, 3, 6])
def test_preshade_and_adjoint(mode, layout, shape):
    rng = np.random.default_rng(100 + mode)
    print(f"preshade mode={mode} layout={layout} shape={shape}", flush=True)
    X, Y, Z = shape
    tf = tf_for(mode, rng)
    d = rng.uniform(-0.1, 1.1, (1, X, Y, Z)).astype(np.float32)
    dt = t32(d) if layout == "contiguous" else x_fastest(d)
    tft = t32(tf[None])
    col = pr.TFUtils.preshade_volume(dt, tft, TF_ENUM[mode])
    torch.cuda.synchronize()
    want = oracle.preshade(mode, tf.astype(np.float32).astype(np.float64), d[0].astype(np.float64))
    got = col.double().cpu().numpy()
    assert got.shape == (4, X, Y, Z)
    if mode == 6:
        err = (np.abs(got - want) / np.maximum(np.abs(want), 1.0)).max()
        assert err < 4e-6, err
    else:
        err = np.abs(got - want).max() / np.abs(want).max()
        assert err < 2e-6, err
    # adjoint through autograd equals the oracle's
    dt.requires_grad_(True)
    adj = rng.standard_normal((4, X, Y, Z)).astype(np.float32)
    col = pr.TFUtils.preshade_volume(dt, tft, TF_ENUM[mode])
    g, = torch.autograd.grad(col, dt, t32(adj))
    gw = oracle.preshade_adjoint(mode, tf.astype(np.float32).astype(np.float64), d[0].astype(np.float64),
                                 adj.astype(np.float64))
    gg = g.double().cpu().numpy()[0]
    if mode in (1, 2):  # FP32 may place a density on the other side of a texel edge / control point
        if mode == 1:
            u = d[0] * tf.shape[0] - 0.5
            safe = np.abs(u - np.round(u)) > 1e-4
        else:
            safe = np.min(np.abs(d[0][..., None] - tf[:, 4].astype(np.float32)), axis=-1) > 1e-5
        safe &= (d[0] > 1e-5) & (d[0] < 1 - 1e-5)
    else:
        safe = np.ones(d[0].shape, bool)
    aerr = np.abs(gg - gw)[safe].max() / np.abs(gw).max()
    print(f"  colour err {err:.1e}, adjoint err {aerr:.1e}")
    assert aerr < 1e-5, aerr


def test_preshade_validation():
    d = torch.zeros(1, 4, 4, 4, device=dev())
    with pytest.raises(RuntimeError, match="unknown tf enum value: TFPreshaded"):
        pr.TFUtils.preshade_volume(d, torch.zeros(1, 2, 6, device=dev()), pr.TFMode.Preshaded)
    with pytest.raises(RuntimeError, match="tf must have 6 entries at dimension 2"):
        pr.TFUtils.preshade_volume(d, torch.zeros(1, 2, 5, device=dev()), pr.TFMode.GaussianLog)
    with pytest.raises(RuntimeError, match="at least two control points"):
        pr.TFUtils.preshade_volume(d, torch.zeros(1, 1, 5, device=dev()), pr.TFMode.Linear)
    with pytest.raises(RuntimeError, match="densityVolume must have 1 entries at dimension 0"):
        pr.TFUtils.preshade_volume(torch.zeros(2, 4, 4, 4, device=dev()), torch.zeros(1, 1, 2, device=dev()),
                                   pr.TFMode.Identity)


@pytest.mark.parametrize("mode,neighbours", [(0, False), (2, False), (3, True), (1, True)])
def test_best_fit_picks_a_minimum_cost_candidate(mode, neighbours):
    rng = np.random.default_rng(200 + mode)
    print(f"best fit mode={mode} neighbours={neighbours}", flush=True)
    X, Y, Z, N = 13, 11, 9, 96
    tf = tf_for(mode, rng).astype(np.float32).astype(np.float64)
    dens = rng.uniform(0, 1, (X, Y, Z))
    target = oracle.preshade(mode, tf, dens) + rng.normal(0, 0.02, (4, X, Y, Z))
    target = target.astype(np.float32).astype(np.float64)
    prev = rng.uniform(0, 1, (X, Y, Z)).astype(np.float32).astype(np.float64) if neighbours else None
    ow, nw = 0.7, 0.25
    torch.manual_seed(5)
    args = (t32(target), t32(tf[None]), TF_ENUM[mode], N, ow)
    if neighbours:
        out = pr.TFUtils.find_best_fit(*args, x_fastest(prev[None]), nw)
    else:
        out = pr.TFUtils.find_best_fit(*args)
    torch.cuda.synchronize()
    torch.manual_seed(5)
    seed = int(torch.randint(0, 2 ** 62, (1,), dtype=torch.int64).item())
    best, best_cost = oracle.best_fit(mode, tf, target, N, ow, prev, nw if neighbours else 0.0, seed)
    got = out.double().cpu().numpy()[0]
    assert out.shape == (1, X, Y, Z)
    same = got == best
    print(f"  equal to the FP64 choice: {same.mean():.4f}")
    assert same.mean() > 0.97
    for v in np.flatnonzero(~same):
        x, y, z = np.unravel_index(v, (X, Y, Z))
        cand = oracle.best_fit_candidates(seed, int(v), N)
        assert got[x, y, z] in cand
        nb, has = None, None
        if neighbours:
            pos = [(x - 1, y, z), (x + 1, y, z), (x, y - 1, z), (x, y + 1, z), (x, y, z - 1), (x, y, z + 1)]
            has = [0 <= p[0] < X and 0 <= p[1] < Y and 0 <= p[2] < Z for p in pos]
            nb = [prev[p] if h else 0.0 for p, h in zip(pos, has)]
        c = oracle.best_fit_cost(mode, tf, target[:, x, y, z], got[x, y, z], ow, nb, has, nw)
        assert c <= best_cost[x, y, z] + 1e-5 * max(1.0, abs(best_cost[x, y, z])), (c, best_cost[x, y, z])


def test_best_fit_validation_and_seeding():
    c = torch.rand(4, 5, 5, 5, device=dev())
    tf = torch.rand(1, 8, 4, device=dev())
    with pytest.raises(RuntimeError, match="numSamples must be >= 1"):
        pr.TFUtils.find_best_fit(c, tf, pr.TFMode.Texture, 0, 1.0)
    with pytest.raises(RuntimeError, match="opacity weighting must be >= 0"):
        pr.TFUtils.find_best_fit(c, tf, pr.TFMode.Texture, 4, -1.0)
    with pytest.raises(RuntimeError, match="unknown tf enum value: TFGaussianLog"):
        pr.TFUtils.find_best_fit(c, torch.rand(1, 2, 6, device=dev()), pr.TFMode.GaussianLog, 4, 1.0)
    with pytest.raises(RuntimeError, match="pd must have 5 entries at dimension 3"):
        pr.TFUtils.find_best_fit(c, tf, pr.TFMode.Texture, 4, 1.0, torch.rand(1, 5, 5, 4, device=dev()), 0.1)
    torch.manual_seed(1)
    a = pr.TFUtils.find_best_fit(c, tf, pr.TFMode.Texture, 16, 1.0)
    torch.manual_seed(1)
    b = pr.TFUtils.find_best_fit(c, tf, pr.TFMode.Texture, 16, 1.0)
    e = pr.TFUtils.find_best_fit(c, tf, pr.TFMode.Texture, 16, 1.0)
    assert torch.equal(a, b) and not torch.equal(a, e)


@pytest.mark.parametrize("mode", [0, 1, 3])
def test_sample_importance(mode):
    rng = np.random.default_rng(300 + mode)
    print(f"importance mode={mode}", flush=True)
    X, Y, Z, N = 21, 17, 19, 50000
    from scipy import ndimage
    d = ndimage.gaussian_filter(rng.uniform(0, 1, (X, Y, Z)), 1.5).astype(np.float32)
    d = (d - d.min()) / (d.max() - d.min())
    tf = tf_for(mode, rng).astype(np.float32).astype(np.float64)
    loc = rng.uniform(0, 1, (N, 3)).astype(np.float32)
    weights = (0.5, 1.0, 1.5, 2.0)
    out = pr.sample_importance(t32(loc), x_fastest(d[None]), t32(tf[None]), TF_ENUM[mode], *weights, 77)
    torch.cuda.synchronize()
    assert out.dtype == torch.bool and out.shape == (N,)
    mx = oracle.importance_maxima(mode, tf, d.astype(np.float64))
    take, margin = oracle.sample_importance(mode, tf, d.astype(np.float64), loc.astype(np.float64), weights, mx, 77)
    got = out.cpu().numpy()
    clear = margin > 1e-5
    print(f"  taken {got.mean():.3f}, mismatches {(got != take).sum()}, near ties {(~clear).sum()}")
    assert np.array_equal(got[clear], take[clear])
    assert (got != take).sum() <= max(3, int((~clear).sum()))


def test_sample_importance_maxima_on_the_device():
    rng = np.random.default_rng(310)
    d = rng.uniform(0, 1, (1, 9, 8, 7)).astype(np.float32)
    tf = np.array([[[1.5, 1.0]]], dtype=np.float32)
    locs = torch.rand(64, 3, device=dev())
    maxima = torch.full((3,), -1.0, device=dev())
    from diffdvr_amd._native import native as n
    out = torch.empty(64, dtype=torch.bool, device=dev())
    n.sample_importance(locs, t32(d), t32(tf), 0, [1.0, 0.0, 0.0, 0.0], 3, maxima, out)
    torch.cuda.synchronize()
    want = oracle.importance_maxima(0, tf[0].astype(np.float64), d[0].astype(np.float64))
    assert np.allclose(maxima.double().cpu().numpy(), want, rtol=1e-5)
    assert bool(out.all())
    with pytest.raises(RuntimeError, match="weights must be non-negative"):
        pr.sample_importance(locs, t32(d), t32(tf), pr.TFMode.Identity, -1, 0, 0, 0, 1)
    with pytest.raises(RuntimeError, match="at least one weight parameter must be positive"):
        pr.sample_importance(locs, t32(d), t32(tf), pr.TFMode.Identity, 0, 0, 0, 0, 1)
    with pytest.raises(RuntimeError, match="sampleLocations must have 3 entries at dimension 1"):
        pr.sample_importance(torch.rand(5, 2, device=dev()), t32(d), t32(tf), pr.TFMode.Identity, 1, 0, 0, 0, 1)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("n,offset", [(1000003, 0), (4099, 1)])
def test_log_losses_and_mul_log(dtype, n, offset):
    rng = np.random.default_rng(400 + n % 7 + offset)
    print(f"losses dtype={dtype} n={n} offset={offset}", flush=True)
    tol = 1e-6 if dtype == torch.float32 else 1e-12
    x = rng.uniform(-3, 3, n + offset)
    y = rng.uniform(-3, 3, n + offset)
    # offset views are 4 or 8 bytes off 16-byte alignment: the scalar path
    tx = torch.tensor(x, dtype=dtype, device=dev())[offset:].requires_grad_(True)
    ty = torch.tensor(y, dtype=dtype, device=dev())[offset:].requires_grad_(True)
    xs, ys = tx.detach().double().cpu().numpy(), ty.detach().double().cpu().numpy()
    g = torch.tensor(rng.standard_normal(n), dtype=dtype, device=dev())
    gs = g.double().cpu().numpy()
    for name, f, fo, fb in (("log_mse", pr.log_mse, oracle.log_mse, oracle.log_mse_backward),
                            ("log_l1", pr.log_l1, oracle.log_l1, oracle.log_l1_backward)):
        out = f(tx, ty)
        want = fo(xs, ys)
        err = np.abs(out.detach().double().cpu().numpy() - want) / np.maximum(np.abs(want), 1.0)
        assert err.max() < 10 * tol, (name, err.max())
        gx, gy = torch.autograd.grad(out, (tx, ty), g)
        wx, wy = fb(xs, ys, gs)
        for got, w in ((gx, wx), (gy, wy)):
            e = np.abs(got.double().cpu().numpy() - w) / np.maximum(np.abs(w), 1.0)
            # the derivative 1 / (1 - e^(y - x)) amplifies the input rounding where x ~ y
            far = np.abs(xs - ys) > 1e-3
            assert e[far].max() < 100 * tol, (name, e[far].max())
    t = torch.tensor(np.concatenate([rng.uniform(0, 2, n - 3), [0.0, 1e-9, 1.0]]), dtype=dtype,
                     device=dev()).requires_grad_(True)
    ts = t.detach().double().cpu().numpy()
    out = pr.mul_log(t)
    assert np.allclose(out.detach().double().cpu().numpy(), oracle.mul_log(ts), rtol=10 * tol, atol=10 * tol)
    gt, = torch.autograd.grad(out, t, g)
    assert np.allclose(gt.double().cpu().numpy(), oracle.mul_log_backward(ts, gs), rtol=10 * tol, atol=10 * tol)


def test_losses_validation():
    a = torch.rand(10, device=dev())
    with pytest.raises(RuntimeError, match="same shape"):
        pr.log_mse(a, torch.rand(11, device=dev()))
    with pytest.raises(RuntimeError, match="same dtype"):
        pr.log_l1(a, torch.rand(10, device=dev(), dtype=torch.float64))
    with pytest.raises(RuntimeError, match="GPU only"):
        pr.mul_log(torch.rand(10))
