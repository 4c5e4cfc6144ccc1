"""The oracle's forward-mode differentiation (dual numbers,
renderer/renderer_kernels.cuh:135-308) pinned on the CPU:

* its value part equals the oracle forward (same march);
* transposition: the Jacobian contracted with a random colour adjoint equals
  both oracle adjoints (exact replay and the reference's inversion), FP64;
* every column equals FP64 central differences of the oracle forward with the
  schedule frozen, on pixels whose difference quotient is smooth;
* the reference's own consistency tests restated
  (unittests/testConsistencyTF.cpp:43-226, testConsistencyCamera.cpp:43-190):
  Marschner-Lobb 32^3, forward-mode gradients contracted with a random
  adjoint equal the adjoint's, symmetric error <= 1e-5 (the reference's EPS),
  in FP64 (both adjoints); the FP32 twin against its inversion adjoint (the
  reference's own pairing) within 1e-4, the FP32 summation error of a TF
  gradient summed over every pixel.
"""
import math

import numpy as np
import pytest
import torch

from helpers import oracle, camera_scene, to_f32, marschner_lobb, consistency_tfs, sym_err
from diffdvr_amd import pyrenderer as pr


def all_indices(s, step=True, camera=True, tf=True):
    D = 0
    kw = {}
    cols = {}
    if step:
        kw["d_step"] = D
        cols["step"] = D
        D += 1
    if camera:
        kw["d_start"] = (D, D + 1, D + 2)
        kw["d_dir"] = (D + 3, D + 4, D + 5)
        cols["ray_start"], cols["ray_dir"] = [D, D + 1, D + 2], [D + 3, D + 4, D + 5]
        D += 6
    if tf:
        _, R, C = s["tf"].shape
        d_tf = np.arange(D, D + R * C, dtype=np.int32).reshape(1, R, C)
        kw["d_tf"] = d_tf
        cols["tf"] = d_tf[0]
        D += R * C
    return D, kw, cols


def contract(G, adj, cols, tf_shape):
    per = np.einsum("bhwdc,bhwc->bhwd", G, adj)
    out = {}
    if "step" in cols:
        out["step"] = per[..., cols["step"]]
    if "ray_start" in cols:
        out["ray_start"] = per[..., cols["ray_start"]]
        out["ray_dir"] = per[..., cols["ray_dir"]]
    if "tf" in cols:
        tot = per.sum(axis=(1, 2))
        tf = np.zeros(tf_shape)
        for b in range(tot.shape[0]):
            tf[b if tf_shape[0] > 1 else 0] += tot[b][cols["tf"]]
        out["tf"] = tf
    return out


def norm_err(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("filter", [0, 1])
@pytest.mark.parametrize("tf_mode", [0, 1, 2, 3])
@pytest.mark.parametrize("blend", [0, 1])
def test_transpose_equals_adjoint(filter, tf_mode, blend):
    rng = np.random.default_rng(7000 + 100 * filter + 10 * tf_mode + blend)
    s = camera_scene(rng, 7, 6, B=2, filter=filter, tf_mode=tf_mode, blend=blend, camera=0, smooth=True,
                     X=12, Y=10, Z=14, step=0.03)
    if blend == 1:
        s["tf"][..., 0 if tf_mode == 0 else 3] *= 0.5
    s["step"] = np.full((1, s["H"], s["W"]), 0.03) + rng.uniform(-0.005, 0.005, (1, s["H"], s["W"]))
    D, kw, cols = all_indices(s)
    c_f, t_f = oracle.forward(s)
    c, t, G = oracle.forward_gradients(s, D, **kw)
    assert np.array_equal(t, t_f)
    assert np.abs(c - c_f).max() <= 1e-13 * max(np.abs(c_f).max(), 1.0)
    adj = rng.standard_normal(c.shape)
    got = contract(G, adj, cols, s["tf"].shape)
    for method in (oracle.EXACT_REPLAY, oracle.INVERSION):
        want = oracle.adjoint(s, c, t, adj, grads=("tf", "camera", "step"), method=method)
        for k, v in want.items():
            if np.linalg.norm(v) == 0:
                assert np.abs(got[k]).max() < 1e-12, (k, method)
                continue
            assert norm_err(got[k], v) < 1e-9, (k, method, norm_err(got[k], v))


@pytest.mark.parametrize("tf_mode", [1, 2, 3])
@pytest.mark.parametrize("blend", [0, 1])
def test_columns_match_central_differences(tf_mode, blend):
    rng = np.random.default_rng(7100 + 10 * tf_mode + blend)
    s = camera_scene(rng, 6, 5, B=1, filter=1, tf_mode=tf_mode, blend=blend, camera=0, smooth=True,
                     X=12, Y=10, Z=14, step=0.03)
    if blend == 1:
        s["tf"][..., 3] *= 0.5
    if tf_mode == 2:  # no clamp active, no density on an end control point
        s["tf"][0, 0, 4], s["tf"][0, -1, 4] = -0.05, 1.05
    D, kw, cols = all_indices(s)
    c, t, t0 = oracle.forward(s, return_t0=True)
    _, _, G = oracle.forward_gradients(s, D, **kw)
    G = G[0]

    def column(key, idx, e):
        sp = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in s.items()}
        sm = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in s.items()}
        sp[key][idx] += e
        sm[key][idx] -= e
        cp, _ = oracle.forward(sp, frozen_term=t, frozen_t0=t0)
        cm, _ = oracle.forward(sm, frozen_term=t, frozen_t0=t0)
        return ((cp - cm) / (2 * e))[0]

    probes = [("step", (Ellipsis,), cols["step"])]
    for k in range(3):
        probes.append(("ray_start", (Ellipsis, k), cols["ray_start"][k]))
        probes.append(("ray_dir", (Ellipsis, k), cols["ray_dir"][k]))
    _, R, C = s["tf"].shape
    for r in range(R):
        for k in range(C):
            probes.append(("tf", (0, r, k), int(cols["tf"][r, k])))
    smooth = []
    for key, idx, d in probes:
        f1, f2 = column(key, idx, 1e-6), column(key, idx, 2.5e-7)
        scale = np.abs(f1).max() + 1e-30
        ok = (np.abs(f1 - f2) <= 1e-7 * scale).all(axis=-1)
        smooth.append(ok.mean())
        # FD rounding is ~1e-16 / 1e-6 of the O(1) colours: absolute floor 1e-9
        assert np.abs(G[:, :, d, :][ok] - f1[ok]).max() <= 1e-6 * scale + 1e-9, (key, idx)
    assert np.mean(smooth) > 0.8


@pytest.fixture
def one_thread():
    """One OpenMP thread: the FP32 adjoint's atomics then sum in a fixed order (reproducible error)."""
    n = oracle.get_threads()
    oracle.set_threads(1)
    yield
    oracle.set_threads(n)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_reference_consistency_tf(dtype, one_thread):
    rng = np.random.default_rng(7200)
    W = H = 48  # the reference renders 128^2; the relation is per pixel
    m = np.array(pr.Camera.compute_matrix(pr.real3(-1, -0.5, 0), pr.real3(0, 0, 0), pr.real3(0, 0, 1), 45, W, H,
                                          0.1, 10.0))
    base = dict(W=W, H=H, filter=1, camera=1, blend=0, early_out=1.0 - 1e-5, fov_y=0.0,
                volume=marschner_lobb(32), box_min=np.full((1, 3), -0.5), box_size=np.ones((1, 3)),
                step=np.full((1, 1, 1), float(np.float32(0.5 / 64))), cam_matrix=to_f32(m[None]))
    for mode, tf in consistency_tfs(rng):
        s = dict(base, tf_mode=mode, tf=tf)
        R, C = tf.shape[1:]
        d_tf = np.arange(R * C, dtype=np.int32).reshape(1, R, C)
        c, t, G = oracle.forward_gradients(s, R * C, d_tf=d_tf, dtype=dtype)
        adj = rng.standard_normal(c.shape).astype(dtype)
        fwd = contract(G.astype(np.float64), adj.astype(np.float64), {"tf": d_tf[0]}, tf.shape)["tf"]
        cf, tt = oracle.forward(s, dtype=dtype)
        assert np.array_equal(t, tt)
        # FP32: the reference's pairing (its forward mode vs its inversion adjoint).  The FP32
        # adjoint sums the TF gradient over all pixels with atomics in thread order, so the
        # comparison carries FP32 summation error (measured 2.3e-5 for the identity TF): 1e-4
        methods = (oracle.INVERSION, oracle.EXACT_REPLAY) if dtype == np.float64 else (oracle.INVERSION,)
        for method in methods:
            want = oracle.adjoint(s, cf, tt, adj, grads=("tf",), dtype=dtype, method=method)["tf"]
            e = sym_err(fwd, want)
            print(f"mode {mode} R={R} method {method} {np.dtype(dtype).name}: symmetric error {e:.2e}")
            assert e <= (1e-5 if dtype == np.float64 else 1e-4), (mode, R, method, e)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_reference_consistency_camera(dtype, one_thread):
    rng = np.random.default_rng(7300)
    W = H = 8
    vp = pr.Camera.viewport_from_lookat(torch.tensor([[-1.0, -0.5, 0.0]]), torch.tensor([[0.0, 0.0, 0.0]]),
                                        torch.tensor([[0.0, 0.0, 1.0]]))
    start, direction = oracle.generate_rays(vp.double().numpy(), H, W, math.radians(45.0))
    for mode, tf in consistency_tfs(rng):
        s = dict(W=W, H=H, filter=1, camera=0, blend=0, early_out=1.0 - 1e-5, fov_y=0.0, tf_mode=mode, tf=tf,
                 volume=marschner_lobb(32), box_min=np.full((1, 3), -0.5), box_size=np.ones((1, 3)),
                 step=np.full((1, 1, 1), 0.25), ray_start=to_f32(start), ray_dir=to_f32(direction))
        c, t, G = oracle.forward_gradients(s, 6, d_dir=(0, 1, 2), d_start=(3, 4, 5), dtype=dtype)
        adj = rng.standard_normal(c.shape).astype(dtype)
        per = np.einsum("bhwdc,bhwc->bhwd", G.astype(np.float64), adj.astype(np.float64))
        cf, tt = oracle.forward(s, dtype=dtype)
        methods = (oracle.INVERSION, oracle.EXACT_REPLAY) if dtype == np.float64 else (oracle.INVERSION,)
        for method in methods:
            want = oracle.adjoint(s, cf, tt, adj, grads=("camera",), dtype=dtype, method=method)
            tol = 1e-5 if dtype == np.float64 else 1e-4  # FP32: the inversion amplifies rounding by 1/(1-A)
            assert sym_err(per[..., 3:6], want["ray_start"]) <= tol, (mode, method)
            assert sym_err(per[..., 0:3], want["ray_dir"]) <= tol, (mode, method)
