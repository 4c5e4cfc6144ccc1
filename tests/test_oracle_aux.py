"""The oracle's auxiliary kernels, pinned without the reference running:
pre-shading (renderer/tf_utils_impl.cu:17-121) against closed forms and central
differences, GaussianLog (renderer/renderer_tf.cuh:653-814) against the log of
the Gaussian TF, the random streams against published sequences, best fit
(:225-520) against a brute-force search, importance sampling
(renderer/losses_impl.cu:254-371) against numpy maxima, trilinear values and
acceptance rates, and the log losses (renderer/renderer_utils.cuh:180-298)
against their closed forms and central differences."""
import numpy as np
import pytest
from scipy import ndimage

from helpers import oracle


def gaussian_tf(rng, R=4):
    tf = np.zeros((R, 6))
    tf[:, :4] = rng.uniform(0.2, 1.5, (R, 4))
    tf[:, 4] = rng.uniform(0.1, 0.9, R)
    tf[:, 5] = rng.uniform(0.08, 0.3, R)
    return tf


def linear_tf(rng, R=5):
    tf = np.zeros((R, 5))
    tf[:, :4] = rng.uniform(0.0, 1.0, (R, 4))
    tf[:, 4] = np.sort(np.concatenate([[0.0, 1.0], rng.uniform(0.1, 0.9, R - 2)]))
    return tf


def tf_for(mode, rng):
    if mode == 0:
        return np.array([[2.5, 0.7]])  # absorption scale, colour scale
    if mode == 1:
        return rng.uniform(0.0, 1.0, (7, 4))
    if mode == 2:
        return linear_tf(rng)
    return gaussian_tf(rng)


def test_random_streams_match_published_sequences():
    # SplitMix64 seeded with 0 (Vigna's reference implementation): first three outputs
    golden = 0x9E3779B97F4A7C15
    assert oracle.splitmix64(0) == 0xE220A8397B1DCDAF
    assert oracle.splitmix64(golden) == 0x6E789E6AA1B965F4
    assert oracle.splitmix64(2 * golden % 2 ** 64) == 0x06C45D188009454F
    # java.util.Random(42).nextInt(): seed scrambled with the multiplier, 32 high bits of 48
    s = (42 ^ 0x5DEECE66D) & (2 ** 48 - 1)
    got = []
    for _ in range(3):
        s = oracle.lcg48_next(s)
        v = (s >> 16) & 0xFFFFFFFF
        got.append(v - 2 ** 32 if v >= 2 ** 31 else v)
    assert got == [-1170105035, 234785527, -1360544799]
    # candidates: 24 high bits of the 48-bit state, in [0, 1), different per voxel
    c0, c1 = oracle.best_fit_candidates(7, 0, 1000), oracle.best_fit_candidates(7, 1, 1000)
    assert np.all((c0 >= 0) & (c0 < 1)) and np.all(c0 * 2 ** 24 == np.floor(c0 * 2 ** 24))
    assert not np.array_equal(c0, c1) and abs(c0.mean() - 0.5) < 0.05


def test_gaussian_log_is_the_log_of_the_gaussian_tf():
    rng = np.random.default_rng(1)
    tf = gaussian_tf(rng, R=6)
    d = rng.uniform(-0.2, 1.2, (5, 4, 3))
    g = oracle.preshade(3, tf, d)
    gl = oracle.preshade(6, tf, d)
    assert np.allclose(np.exp(gl), g, rtol=1e-12, atol=0)
    # far from every mean the sum underflows in linear space, not in log space
    tf1 = tf[:1].copy()
    tf1[0, 4], tf1[0, 5] = 0.0, 0.01
    gl1 = oracle.preshade(6, tf1, np.array([[[1.0]]]))
    assert np.allclose(gl1[:, 0, 0, 0], np.log(tf1[0, :4]) - 1.0 / (2 * 0.01 ** 2), rtol=1e-14)


def test_preshade_closed_forms():
    rng = np.random.default_rng(2)
    d = rng.uniform(-0.3, 1.3, (4, 3, 5))
    # identity: clamp(d) * (c, c, c, a)
    col = oracle.preshade(0, [[2.5, 0.7]], d)
    dc = np.clip(d, 0, 1)
    assert np.allclose(col, np.stack([0.7 * dc, 0.7 * dc, 0.7 * dc, 2.5 * dc]), rtol=0, atol=1e-15)
    # texture, R = 7: texel centres (i + 0.5) / R give row i
    tex = rng.uniform(0, 1, (7, 4))
    centres = ((np.arange(7) + 0.5) / 7).reshape(7, 1, 1)
    assert np.allclose(oracle.preshade(1, tex, centres)[:, :, 0, 0].T, tex, atol=1e-14)
    # linear: control points give their rows, midpoints the mean of neighbouring rows
    lin = linear_tf(rng)
    pos = lin[:, 4].reshape(-1, 1, 1)
    assert np.allclose(oracle.preshade(2, lin, pos)[:, :, 0, 0].T, lin[:, :4], atol=1e-14)
    mid = ((lin[1:, 4] + lin[:-1, 4]) / 2).reshape(-1, 1, 1)
    assert np.allclose(oracle.preshade(2, lin, mid)[:, :, 0, 0].T, (lin[1:, :4] + lin[:-1, :4]) / 2, atol=1e-14)
    # one Gaussian at its mean gives its colour
    g = gaussian_tf(rng, R=1)
    assert np.allclose(oracle.preshade(3, g, np.full((1, 1, 1), g[0, 4]))[:, 0, 0, 0], g[0, :4], rtol=1e-15)


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 6])
def test_preshade_adjoint_central_differences(mode):
    rng = np.random.default_rng(10 + mode)
    tf = tf_for(mode, rng)
    d = rng.uniform(0.05, 0.95, (4, 3, 5))
    if mode == 1:  # keep away from texel edges (d * R - 0.5 integer)
        u = d * 7 - 0.5
        d = np.where(np.abs(u - np.round(u)) < 0.02, d + 0.01, d)
    if mode == 2:  # and from control points
        for p in tf[:, 4]:
            d = np.where(np.abs(d - p) < 0.01, d + 0.02, d)
    adj = rng.standard_normal((4,) + d.shape)
    g = oracle.preshade_adjoint(mode, tf, d, adj)
    h = 1e-6
    fd = (np.sum(adj * oracle.preshade(mode, tf, d + h), axis=0) -
          np.sum(adj * oracle.preshade(mode, tf, d - h), axis=0)) / (2 * h)
    assert np.allclose(g, fd, rtol=1e-6, atol=1e-7), np.abs(g - fd).max()


def test_preshade_adjoint_identity_mask():
    """Outside (0, 1) the identity TF is clamped and the density gradient is 0 (renderer_tf.cuh:103)."""
    d = np.array([[[-0.1, 0.0, 0.5, 1.0, 1.2]]])
    g = oracle.preshade_adjoint(0, [[2.0, 3.0]], d, np.ones((4,) + d.shape))
    assert np.array_equal(g[0, 0], [0.0, 0.0, 3 * 3.0 + 2.0, 0.0, 0.0])


def numpy_identity_colour(d, sa, sc):
    dc = np.clip(d, 0, 1)
    return np.stack([sc * dc, sc * dc, sc * dc, sa * dc], axis=-1)


@pytest.mark.parametrize("neighbours", [False, True])
def test_best_fit_is_the_brute_force_minimum(neighbours):
    rng = np.random.default_rng(20 + neighbours)
    X, Y, Z, N = 3, 4, 2, 50
    sa, sc = 1.7, 0.6
    target = rng.uniform(0, 1, (4, X, Y, Z))
    prev = rng.uniform(0, 1, (X, Y, Z)) if neighbours else None
    ow, nw, seed = 0.8, 0.3, 12345
    out, cost = oracle.best_fit(0, [[sa, sc]], target, N, ow, prev, nw, seed)
    for v in range(X * Y * Z):
        x, y, z = v // (Y * Z), (v // Z) % Y, v % Z
        cand = oracle.best_fit_candidates(seed, v, N)
        col = numpy_identity_colour(cand, sa, sc)
        t = target[:, x, y, z]
        c = np.sum((col[:, :3] - t[:3]) ** 2, axis=1) + ow * np.log(1 + np.abs(col[:, 3] - t[3]))
        if neighbours:
            for dx, dy, dz in ((-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)):
                q = (x + dx, y + dy, z + dz)
                if 0 <= q[0] < X and 0 <= q[1] < Y and 0 <= q[2] < Z:
                    c = c + nw * (cand - prev[q]) ** 2
        k = int(np.argmin(c))
        assert out[x, y, z] == cand[k] and cost[x, y, z] == pytest.approx(c[k], rel=1e-12)
        assert oracle.best_fit_cost(0, [[sa, sc]], t, cand[k], ow) == pytest.approx(
            np.sum((col[k, :3] - t[:3]) ** 2) + ow * np.log1p(abs(col[k, 3] - t[3])), rel=1e-12)


def test_best_fit_recovers_densities_through_an_invertible_tf():
    rng = np.random.default_rng(22)
    lin = np.array([[0.0, 0.1, 0.9, 0.0, 0.0], [0.5, 0.5, 0.2, 1.0, 0.5], [1.0, 0.9, 0.0, 3.0, 1.0]])
    dens = rng.uniform(0.02, 0.98, (5, 4, 3))
    colour = oracle.preshade(2, lin, dens)
    out, cost = oracle.best_fit(2, lin, colour, 4096, 1.0, seed=99)
    # the nearest of 4096 uniform candidates lies within ~1e-3 of the true density; the
    # opacity term log(1 + |a - a_tf|) is linear in that distance (slope <= 4 here)
    assert np.abs(out - dens).max() < 5e-3 and cost.max() < 2e-2


def trilinear_numpy(vol, p):
    return ndimage.map_coordinates(vol, p.T, order=1, mode="nearest")


def test_importance_maxima_and_trilinear_values():
    rng = np.random.default_rng(30)
    X, Y, Z = 6, 5, 7
    d = rng.uniform(0, 1, (X, Y, Z))
    sa = 2.0
    m = oracle.importance_maxima(0, [[sa, 1.0]], d)
    pad = np.pad(d, 1, mode="edge")
    gx = (pad[2:, 1:-1, 1:-1] - pad[:-2, 1:-1, 1:-1]) / 2
    gy = (pad[1:-1, 2:, 1:-1] - pad[1:-1, :-2, 1:-1]) / 2
    gz = (pad[1:-1, 1:-1, 2:] - pad[1:-1, 1:-1, :-2]) / 2
    assert m[0] == pytest.approx(np.sqrt(gx ** 2 + gy ** 2 + gz ** 2).max(), rel=1e-14)
    assert m[1] == pytest.approx(sa * d.max(), rel=1e-14)
    assert m[2] == pytest.approx(sa * np.sqrt(gx ** 2 + gy ** 2 + gz ** 2).max(), rel=1e-14)
    # the sampler's trilinear fetch agrees with scipy inside the volume
    p = rng.uniform(0, 1, (50, 3)) * [X - 1, Y - 1, Z - 1]
    lib_vals = np.array([oracle.fetch(1, d, q) for q in p])
    assert np.allclose(lib_vals, trilinear_numpy(d, p), atol=1e-14)


def test_importance_decisions_follow_their_probabilities():
    rng = np.random.default_rng(31)
    X, Y, Z = 9, 8, 7
    d = ndimage.gaussian_filter(rng.uniform(0, 1, (X, Y, Z)), 1.0)
    tf = [[1.0, 1.0]]
    N = 20000
    loc = rng.uniform(0, 1, (N, 3))
    mx = oracle.importance_maxima(0, tf, d)
    # uniform only: every sample is taken
    take, margin = oracle.sample_importance(0, tf, d, loc, (1, 0, 0, 0), mx, 5)
    assert take.all() and np.isinf(margin).all()
    # opacity only: taken with probability opacity / max opacity
    take, _ = oracle.sample_importance(0, tf, d, loc, (0, 0, 3, 0), mx, 5)
    p = np.clip(trilinear_numpy(d, loc * [X - 1, Y - 1, Z - 1]), 0, 1) / mx[1]
    assert abs(take.mean() - p.mean()) < 4 * np.sqrt(p.mean() * (1 - p.mean()) / N)
    # half uniform, half opacity
    take, margin = oracle.sample_importance(0, tf, d, loc, (1, 0, 1, 0), mx, 6)
    want = 0.5 + 0.5 * p.mean()
    assert abs(take.mean() - want) < 4 * np.sqrt(want * (1 - want) / N)
    assert 0.45 < np.isinf(margin).mean() < 0.55
    # thresholds restate the reference's float normalisation
    assert np.array_equal(oracle.importance_thresholds((1, 2, 3, 4)), np.float32([0.1, 0.3, 0.6]))


def test_log_losses_closed_forms_and_derivatives():
    rng = np.random.default_rng(40)
    x, y = rng.uniform(0.01, 5, 1000), rng.uniform(0.01, 5, 1000)
    lx, ly = np.log(x), np.log(y)
    assert np.allclose(np.exp(oracle.log_mse(lx, ly)), (x - y) ** 2, rtol=1e-10)
    assert np.allclose(np.exp(oracle.log_l1(lx, ly)), np.abs(x - y), rtol=1e-10)
    # nearly equal arguments keep their relative accuracy
    assert oracle.log_l1(np.log(1.0), np.log1p(1e-12)) == pytest.approx(np.log(1e-12), rel=1e-6)
    g = rng.standard_normal(1000)
    h = 1e-6
    far = np.abs(lx - ly) > 0.05  # central differences need the kink at x = y out of reach
    lx, ly, g = lx[far], ly[far], g[far]
    for f, fb in ((oracle.log_mse, oracle.log_mse_backward), (oracle.log_l1, oracle.log_l1_backward)):
        gx, gy = fb(lx, ly, g)
        assert np.allclose(gx, g * (f(lx + h, ly) - f(lx - h, ly)) / (2 * h), rtol=1e-5, atol=1e-6)
        assert np.allclose(gy, g * (f(lx, ly + h) - f(lx, ly - h)) / (2 * h), rtol=1e-5, atol=1e-6)
    t = np.concatenate([rng.uniform(0, 1, 100), [0.0, 1e-8, 1.0]])
    ref = np.where(t < 1.192092896e-07, 0.0, t * np.log(np.maximum(t, 1e-300)) / np.log(2))
    assert np.allclose(oracle.mul_log(t), ref, rtol=1e-14, atol=0)
    tt = rng.uniform(0.01, 1, 100)
    gb = oracle.mul_log_backward(tt, np.ones_like(tt))
    assert np.allclose(gb, (oracle.mul_log(tt + h) - oracle.mul_log(tt - h)) / (2 * h), rtol=1e-7)
