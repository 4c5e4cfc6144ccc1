"""Pins the oracle's components with the reference's own component tests,
restated with central differences in FP64 (unittests/check_adjoint.h:35-84).

  blending       unittests/testAdjointBlending.cpp:10-78
  TF             unittests/testAdjointTF.cpp:100-233
  interpolation  unittests/testAdjointInterpolation.cpp:23-108
  cross/normalize unittests/testAdjointUtils.cpp:11-88
plus closed forms a dropped term, wrong sign or swapped index would fail.
"""
import math

import numpy as np
import pytest

from helpers import oracle


def check_adjoint(x, forward, adjoint, eps=1e-6, rtol=1e-6, atol=1e-9, trials=None, rng=None):
    """For unit and random output seeds g: adjoint(x, g) == grad_x <g, forward(x)> by central FD."""
    rng = rng or np.random.default_rng(42)
    y = forward(x)
    m, n = y.size, x.size
    seeds = [np.eye(m)[i] for i in range(m)] + [rng.uniform(-1, 1, m) for _ in range(trials or 2)]
    fd = np.zeros((n, m))
    for i in range(n):
        xp, xm = x.copy(), x.copy()
        xp[i] += eps
        xm[i] -= eps
        fd[i] = (forward(xp) - forward(xm)) / (2 * eps)
    for g in seeds:
        got = adjoint(x, g)
        want = fd @ g
        np.testing.assert_allclose(got, want, rtol=rtol, atol=atol)


@pytest.mark.parametrize("mode", [0, 1])
def test_blend_adjoint_fd(mode):
    rng = np.random.default_rng(42)
    for _ in range(20):
        x = rng.uniform(0.01, 0.99, 9)

        def fwd(v):
            return oracle.blend(mode, v[0:4], v[4:8], v[8])

        def adj(v, g):
            out = oracle.blend(mode, v[0:4], v[4:8], v[8])
            acc_in, adj_acc, adj_cur, adj_step = oracle.blend_adjoint(mode, out, v[4:8], v[8], g)
            # the inversion reconstructs the accumulator (testAdjointBlending.cpp:44-49)
            np.testing.assert_allclose(acc_in, v[0:4], rtol=1e-9, atol=1e-12)
            return np.concatenate([adj_acc, adj_cur, [adj_step]])

        check_adjoint(x, fwd, adj, eps=1e-7, rtol=1e-6, atol=1e-9)


def test_blend_closed_forms():
    acc = np.array([0.1, 0.2, 0.3, 0.4])
    cur = np.array([0.5, 0.6, 0.7, 2.0])
    out = oracle.blend(0, acc, cur, 0.25)
    a = 1 - math.exp(-0.5)
    np.testing.assert_allclose(out, acc + (1 - 0.4) * np.array([0.5 * a, 0.6 * a, 0.7 * a, a]), rtol=1e-14)
    out = oracle.blend(1, acc, cur, 0.25)  # alpha = min(1, 2*0.25) = 0.5
    np.testing.assert_allclose(out, acc + 0.6 * np.array([0.25, 0.3, 0.35, 0.5]), rtol=1e-14)
    out = oracle.blend(1, acc, cur, 0.75)  # alpha clamps to 1: accumulator becomes opaque
    assert out[3] == pytest.approx(1.0)


def make_tf(mode, rng):
    if mode == 0:
        return rng.uniform(0, 1, (1, 2))
    if mode == 1:
        return rng.uniform(0, 1, (8, 4))
    if mode == 2:
        pos = np.concatenate([[0.0], np.sort(rng.uniform(0.1, 0.9, 3)), [1.0]])
        return np.concatenate([rng.uniform(0.1, 0.9, (5, 4)), pos[:, None]], axis=1)
    tf = np.zeros((4, 6))
    tf[:, :4] = rng.uniform(0, 1, (4, 4))
    tf[:, 4] = rng.uniform(0, 1, 4)
    tf[:, 5] = 0.1 + rng.uniform(0, 1, 4)
    return tf


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
@pytest.mark.parametrize("with_tf", [False, True])
def test_tf_adjoint_fd(mode, with_tf):
    rng = np.random.default_rng(42)
    for _ in range(6):
        tf = make_tf(mode, rng)
        for _ in range(6):
            d = rng.uniform(0.01, 0.99)
            x = np.concatenate([[d], tf.ravel()]) if with_tf else np.array([d])

            def fwd(v):
                t = v[1:].reshape(tf.shape) if with_tf else tf
                return oracle.tf_eval(mode, t, v[0])

            def adj(v, g):
                t = v[1:].reshape(tf.shape) if with_tf else tf
                ad, gt = oracle.tf_adjoint(mode, t, v[0], g, want_tf=with_tf)
                return np.concatenate([[ad], gt.ravel()]) if with_tf else np.array([ad])

            check_adjoint(x, fwd, adj, eps=1e-7, rtol=1e-6, atol=1e-8)


def test_tf_closed_forms():
    rng = np.random.default_rng(7)
    # texture: exact at texel centres d=(i+0.5)/R, clamped beyond the ends
    tex = rng.uniform(0, 1, (8, 4))
    for i in range(8):
        np.testing.assert_allclose(oracle.tf_eval(1, tex, (i + 0.5) / 8), tex[i], rtol=1e-12)
    np.testing.assert_allclose(oracle.tf_eval(1, tex, 0.0), tex[0], rtol=1e-12)
    np.testing.assert_allclose(oracle.tf_eval(1, tex, 1.0), tex[7], rtol=1e-12)
    np.testing.assert_allclose(oracle.tf_eval(1, tex, 1.5 / 8 + 0.25 / 8), 0.75 * tex[1] + 0.25 * tex[2], rtol=1e-12)
    # linear: exact at control points, midpoint average
    lin = make_tf(2, rng)
    for i in range(1, 4):
        np.testing.assert_allclose(oracle.tf_eval(2, lin, lin[i, 4]), lin[i, :4], rtol=1e-12)
    mid = 0.5 * (lin[1, 4] + lin[2, 4])
    np.testing.assert_allclose(oracle.tf_eval(2, lin, mid), 0.5 * (lin[1, :4] + lin[2, :4]), rtol=1e-12)
    # identity: clamp to [0,1], (d*sc, d*sc, d*sc, d*sa)
    np.testing.assert_allclose(oracle.tf_eval(0, np.array([[3.0, 2.0]]), 0.25), [0.5, 0.5, 0.5, 0.75])
    np.testing.assert_allclose(oracle.tf_eval(0, np.array([[3.0, 2.0]]), 1.7), [2.0, 2.0, 2.0, 3.0])
    # identity adjoint is zero outside (0,1)
    assert oracle.tf_adjoint(0, np.array([[3.0, 2.0]]), 1.2, np.ones(4), want_tf=False)[0] == 0.0
    # gaussian: column 5 is sigma: c at mu, c*exp(-1/2) at mu+sigma
    g = np.array([[0.2, 0.4, 0.6, 0.8, 0.3, 0.05]])
    np.testing.assert_allclose(oracle.tf_eval(3, g, 0.3), g[0, :4], rtol=1e-14)
    np.testing.assert_allclose(oracle.tf_eval(3, g, 0.35), g[0, :4] * math.exp(-0.5), rtol=1e-12)


@pytest.mark.parametrize("with_volume", [False, True])
def test_trilinear_adjoint_fd(with_volume):
    rng = np.random.default_rng(42)
    res = (2, 3, 4)
    base = rng.uniform(0.01, 0.99, res)
    for _ in range(20):
        p = rng.uniform(0.01, 0.99, 3) * (np.array(res) - 1)
        x = np.concatenate([p, base.ravel()]) if with_volume else p

        def fwd(v):
            vol = v[3:].reshape(res) if with_volume else base
            return np.array([oracle.fetch(1, vol, v[:3])])

        def adj(v, g):
            vol = v[3:].reshape(res) if with_volume else base
            ap, gv = oracle.trilinear_adjoint(vol, v[:3], g[0], want_volume=with_volume)
            return np.concatenate([ap, gv.ravel()]) if with_volume else ap

        check_adjoint(x, fwd, adj, eps=1e-7, rtol=1e-6, atol=1e-9)


def test_interpolation_closed_forms():
    rng = np.random.default_rng(3)
    X, Y, Z = 5, 6, 7
    a, b, c, d = 0.3, -0.7, 1.1, 0.25
    ii, jj, kk = np.meshgrid(np.arange(X), np.arange(Y), np.arange(Z), indexing="ij")
    vol = a * ii + b * jj + c * kk + d
    for _ in range(10):
        p = rng.uniform(0, [X - 1, Y - 1, Z - 1])
        assert oracle.fetch(1, vol, p) == pytest.approx(a * p[0] + b * p[1] + c * p[2] + d, rel=1e-12)
        ap, _ = oracle.trilinear_adjoint(vol, p, 2.0, want_volume=False)
        np.testing.assert_allclose(ap, 2.0 * np.array([a, b, c]), rtol=1e-12)
    # corner weights sum to 1 and reproduce the interpolation weights
    p = np.array([1.25, 2.5, 3.75])
    _, gv = oracle.trilinear_adjoint(np.zeros((X, Y, Z)), p, 1.0)
    assert gv.sum() == pytest.approx(1.0)
    assert gv[1, 2, 3] == pytest.approx(0.75 * 0.5 * 0.25)
    assert gv[2, 3, 4] == pytest.approx(0.25 * 0.5 * 0.75)
    # truncating cell index + unclamped fraction: slightly negative positions extrapolate
    assert oracle.fetch(1, vol, np.array([-0.25, 1.0, 1.0])) == pytest.approx(a * -0.25 + b + c + d, rel=1e-12)
    # nearest: round half away from zero, clamp into the volume
    v = rng.uniform(0, 1, (X, Y, Z))
    assert oracle.fetch(0, v, np.array([1.5, 2.49, 3.51])) == v[2, 2, 4]
    assert oracle.fetch(0, v, np.array([-3.0, 9.0, 2.2])) == v[0, Y - 1, 2]


def test_adj_cross_fd():
    rng = np.random.default_rng(42)
    for _ in range(20):
        x = rng.standard_normal(6)
        check_adjoint(x, lambda v: np.cross(v[:3], v[3:]),
                      lambda v, g: np.concatenate(oracle.adj_cross(v[:3], v[3:], g)), eps=1e-6, rtol=1e-7)


def test_adj_normalize_fd():
    rng = np.random.default_rng(42)
    for _ in range(20):
        x = rng.standard_normal(3)
        check_adjoint(x, lambda v: v / np.linalg.norm(v), lambda v, g: oracle.adj_normalize(v, g),
                      eps=1e-6, rtol=1e-6, atol=1e-9)


def test_fp32_twin_tracks_fp64():
    rng = np.random.default_rng(5)
    for mode in range(4):
        tf = make_tf(mode, rng)
        for _ in range(5):
            d = rng.uniform(0.01, 0.99)
            np.testing.assert_allclose(oracle.tf_eval(mode, tf, d, dtype=np.float32), oracle.tf_eval(mode, tf, d),
                                       rtol=1e-5, atol=1e-6)
