"""Caller-side parametrizations of diffdvr (pytests/diffdvr/parametrizations.py, renderer.py:53-58):
pinned by closed forms, inverse-function round trips and a hand-derived peak table (CPU)."""
import math

import numpy as np
import pytest
import torch
import torch.nn.functional as F

from diffdvr_amd import diffdvr as dd
from diffdvr_amd import pyrenderer as pr


def test_inverse_sigmoid_closed_forms_and_round_trip():
    assert float(dd.inverseSigmoid(torch.tensor(0.5, dtype=torch.float64))) == 0.0
    y = torch.tensor(1.0 / (1.0 + math.e), dtype=torch.float64)          # sigmoid(-1)
    assert abs(float(dd.inverseSigmoid(y)) + 1.0) < 1e-15
    x = torch.linspace(-12, 12, 97, dtype=torch.float64)
    assert torch.allclose(dd.inverseSigmoid(torch.sigmoid(x)), x, rtol=0, atol=1e-9)
    assert float(dd.inverseSigmoid(torch.tensor(0.0))) == -math.inf


@pytest.mark.parametrize("beta", [1.0, 2.0, 0.5])
def test_inverse_softplus_round_trip_and_threshold(beta):
    # log(exp(beta*y) - 1) / beta at y = log(2) / beta is 0
    assert abs(float(dd.inverseSoftplus(torch.tensor(math.log(2.0) / beta, dtype=torch.float64), beta))) < 1e-15
    x = torch.linspace(-6, 8, 57, dtype=torch.float64) / beta
    assert torch.allclose(dd.inverseSoftplus(F.softplus(x, beta, 20), beta, 20), x, rtol=0, atol=1e-9)
    # above the threshold softplus is the identity, and so is its inverse
    big = torch.tensor([25.0 / beta, 40.0 / beta], dtype=torch.float64)
    assert torch.equal(dd.inverseSoftplus(big, beta, 20), big)
    assert torch.equal(F.softplus(big, beta, 20), big)


def test_volume_densities_round_trip_and_gradient():
    rng = np.random.default_rng(3)
    v = torch.from_numpy(rng.uniform(0.01, 0.99, size=(4, 5, 6)))
    m = dd.VolumeDensities()
    x = dd.VolumeDensities.prepare_input(v).requires_grad_(True)
    out = m(x)
    assert torch.allclose(out, v, rtol=0, atol=1e-12)
    w = torch.from_numpy(rng.normal(size=(4, 5, 6)))
    (out * w).sum().backward()
    assert torch.allclose(x.grad, w * v * (1 - v), rtol=1e-12, atol=0)      # sigmoid' = s (1 - s)


def test_tf_texture_init_from_points_inverts_forward():
    pts = [pr.TFPoint(0.0, pr.real4(0.1, 0.2, 0.3, 0.0)), pr.TFPoint(0.4, pr.real4(0.9, 0.5, 0.05, 3.0)),
           pr.TFPoint(1.0, pr.real4(0.3, 0.7, 0.6, 25.0))]
    R = 64
    tex = pr.TFUtils.get_texture_tensor(pts, R)
    pre = dd.TfTexture.init_from_points(pts, R)
    assert pre.shape == (1, R, 4)
    back = dd.TfTexture()(pre)
    assert torch.allclose(back, tex, rtol=1e-5, atol=1e-6)
    # the last texel's opacity (~24.8) is above the softplus threshold: its own parameter
    assert float(tex[0, R - 1, 3]) > 20 and float(pre[0, R - 1, 3]) == float(tex[0, R - 1, 3])


def test_tf_texture_ranges():
    x = torch.from_numpy(np.random.default_rng(5).normal(scale=8, size=(2, 16, 4)))
    y = dd.TfTexture()(x)
    assert float(y[..., :3].min()) > 0 and float(y[..., :3].max()) < 1
    assert float(y[..., 3].min()) >= 0
    assert torch.allclose(y[..., :3], torch.sigmoid(x[..., :3]))
    assert torch.allclose(y[..., 3:], F.softplus(x[..., 3:]))


def test_piecewise_linear_from_peaks_hand_derived():
    # peak A at 0.3 (width 0.1), peak B at 0.7 (width 0.2), given in reverse order
    peaks = np.array([[0, 1, 0, 4, 0.7, 0.2], [1, 0, 0, 2, 0.3, 0.1]], dtype=np.float64)
    tf = dd.TfPiecewiseLinear.create_from_peaks(peaks, opacity_scaling=0.5)
    want = np.array([[0, 0, 0, 0, 0.0],
                     [1, 0, 0, 0, 0.2], [1, 0, 0, 1.0, 0.3], [1, 0, 0, 0, 0.4],     # A: [0.3 - 0.1, min(0.5, 0.4)]
                     [0, 1, 0, 0, 0.5], [0, 1, 0, 2.0, 0.7], [0, 1, 0, 0, 0.9],     # B: [max(0.4, 0.5), 0.7 + 0.2]
                     [0, 0, 0, 0, 1.0]], dtype=np.float32)[None]
    assert tf.dtype == torch.float32 and tf.shape == (1, 8, 5)
    assert np.allclose(tf.numpy(), want, atol=1e-7)


def test_piecewise_linear_positions_sorted_and_bounded():
    rng = np.random.default_rng(11)
    for _ in range(50):
        n = int(rng.integers(1, 6))
        peaks = np.concatenate([rng.uniform(0, 1, (n, 4)), rng.uniform(0, 1, (n, 1)), rng.uniform(0.01, 0.4, (n, 1))],
                               axis=1)
        tf = dd.TfPiecewiseLinear.create_from_peaks(peaks)[0].numpy()
        pos = tf[:, 4]
        assert pos[0] == 0.0 and pos[-1] == 1.0
        assert np.all(np.diff(pos) >= -1e-7), pos
        # every peak's opacity appears at its position
        for p in peaks:
            hit = np.where(np.abs(pos - np.float32(p[4])) < 1e-7)[0]
            assert any(abs(tf[h, 3] - np.float32(p[3])) < 1e-6 for h in hit)


def test_camera_on_a_sphere_random_points():
    np.random.seed(123)
    N = 20000
    center, yaw, pitch, dist = dd.CameraOnASphere.random_points(N, min_zoom=6, max_zoom=7, center=[0.1, 0.2, 0.3])
    assert center.shape == (N, 3) and yaw.shape == (N, 1) and pitch.shape == (N, 1) and dist.shape == (N, 1)
    assert center.dtype == torch.float32
    assert torch.equal(center, torch.tensor([[0.1, 0.2, 0.3]] * N))
    p, y = pitch.double().numpy()[:, 0], yaw.double().numpy()[:, 0]
    assert p.min() >= 0 and p.max() <= math.pi and y.min() >= -math.pi and y.max() <= math.pi
    d = dist.double().numpy()[:, 0]
    assert d.min() >= 1.1 ** 6 - 1e-5 and d.max() <= 1.1 ** 7 + 1e-5
    # uniform on the sphere: z = cos(pitch) has mean 0 and second moment 1/3; x, y likewise
    z = np.cos(p)
    xs, ys = np.sin(p) * np.cos(y), np.sin(p) * np.sin(y)
    for c in (xs, ys, z):
        assert abs(c.mean()) < 0.02 and abs((c * c).mean() - 1.0 / 3.0) < 0.02
    # same global generator state, same points
    np.random.seed(123)
    c2, y2, p2, d2 = dd.CameraOnASphere.random_points(N, min_zoom=6, max_zoom=7, center=[0.1, 0.2, 0.3])
    assert torch.equal(y2, yaw) and torch.equal(p2, pitch) and torch.equal(d2, dist)
