"""ctypes front end of the CPU oracle (liboracle_dvr.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by the cpu_baseline leg of bench.py, always as the checker or the timed CPU
baseline, never as part of the product.

A *scene* is a dict of numpy arrays and modes (all batched arrays carry their
own leading batch size, 1 = broadcast):
    W, H, filter, camera, tf_mode, blend, early_out, fov_y
    volume (vb,X,Y,Z), box_min (nb,3), box_size (nb,3), step (sb,sh,sw), tf (tb,R,C)
    camera 0: ray_start, ray_dir (rb,H,W,3); 1: cam_matrix (cb,4,4); 2: cam_matrix (cb,3,3)
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_dvr.so")

INVERSION = 0
EXACT_REPLAY = 1


class _Inputs(ctypes.Structure):
    _fields_ = [
        ("B", ctypes.c_int), ("W", ctypes.c_int), ("H", ctypes.c_int),
        ("filter", ctypes.c_int), ("camera", ctypes.c_int), ("tf_mode", ctypes.c_int), ("blend", ctypes.c_int),
        ("volume", ctypes.c_void_p), ("vol_b", ctypes.c_int), ("X", ctypes.c_int), ("Y", ctypes.c_int),
        ("Z", ctypes.c_int),
        ("box_min", ctypes.c_void_p), ("box_min_b", ctypes.c_int),
        ("box_size", ctypes.c_void_p), ("box_size_b", ctypes.c_int),
        ("ray_start", ctypes.c_void_p), ("ray_dir", ctypes.c_void_p), ("ray_b", ctypes.c_int),
        ("cam_matrix", ctypes.c_void_p), ("cam_b", ctypes.c_int),
        ("fov_y", ctypes.c_double),
        ("step", ctypes.c_void_p), ("step_b", ctypes.c_int), ("step_h", ctypes.c_int), ("step_w", ctypes.c_int),
        ("tf", ctypes.c_void_p), ("tf_b", ctypes.c_int), ("R", ctypes.c_int), ("C", ctypes.c_int),
        ("early_out", ctypes.c_double),
    ]


class _AdjOut(ctypes.Structure):
    _fields_ = [
        ("adj_volume", ctypes.c_void_p), ("adj_volume_b", ctypes.c_int),
        ("adj_tf", ctypes.c_void_p), ("adj_tf_b", ctypes.c_int),
        ("adj_ray_start", ctypes.c_void_p), ("adj_ray_dir", ctypes.c_void_p), ("adj_cam_b", ctypes.c_int),
        ("adj_step", ctypes.c_void_p), ("adj_step_b", ctypes.c_int), ("adj_step_h", ctypes.c_int),
        ("adj_step_w", ctypes.c_int),
    ]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        for suf in ("f64", "f32"):
            getattr(_lib, f"oracle_tf_adjoint_{suf}").restype = ctypes.c_double if suf == "f64" else ctypes.c_float
            getattr(_lib, f"oracle_fetch_{suf}").restype = ctypes.c_double if suf == "f64" else ctypes.c_float
        _lib.oracle_get_threads.restype = ctypes.c_int
        for suf, ct in (("f64", ctypes.c_double), ("f32", ctypes.c_float)):
            getattr(_lib, f"oracle_best_fit_cost_{suf}").restype = ct
        for name in ("oracle_splitmix64", "oracle_lcg48_next", "oracle_best_fit_stream"):
            getattr(_lib, name).restype = ctypes.c_uint64
        _lib.oracle_splitmix64.argtypes = [ctypes.c_uint64]
        _lib.oracle_lcg48_next.argtypes = [ctypes.c_uint64]
        _lib.oracle_best_fit_stream.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
    return _lib


def set_threads(n: int):
    lib().oracle_set_threads(int(n))


def get_threads() -> int:
    return lib().oracle_get_threads()


def _suffix(dtype):
    return "f64" if np.dtype(dtype) == np.float64 else "f32"


def _c(dtype):
    return ctypes.c_double if np.dtype(dtype) == np.float64 else ctypes.c_float


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def batch_size(scene) -> int:
    B = 1
    # a pre-shaded volume (filter 3) holds colour channels in its batch dimension
    keys = (["volume"] if int(scene["filter"]) != 3 else []) + ["box_min", "box_size", "step", "tf"]
    keys += ["ray_start", "ray_dir"] if scene["camera"] == 0 else ["cam_matrix"]
    for k in keys:
        b = scene[k].shape[0]
        if b != 1:
            if B not in (1, b):
                raise ValueError(f"batch mismatch at {k}")
            B = b
    return B


class _Prepared:
    """Contiguous copies of the scene arrays in one dtype, kept alive with the struct."""

    def __init__(self, scene, dtype):
        self.dtype = np.dtype(dtype)
        conv = lambda a: None if a is None else np.ascontiguousarray(np.asarray(a, dtype=self.dtype))
        s = {k: conv(scene.get(k)) for k in ("volume", "box_min", "box_size", "ray_start", "ray_dir", "cam_matrix",
                                             "step", "tf")}
        self.arrays = s
        self.B = batch_size(scene)
        W, H = int(scene["W"]), int(scene["H"])
        st = s["step"]
        st3 = st.reshape(st.shape[0], st.shape[1], st.shape[2])
        self.s = _Inputs(
            B=self.B, W=W, H=H, filter=int(scene["filter"]), camera=int(scene["camera"]),
            tf_mode=int(scene["tf_mode"]), blend=int(scene["blend"]),
            volume=_ptr(s["volume"]), vol_b=s["volume"].shape[0], X=s["volume"].shape[1], Y=s["volume"].shape[2],
            Z=s["volume"].shape[3],
            box_min=_ptr(s["box_min"]), box_min_b=s["box_min"].shape[0],
            box_size=_ptr(s["box_size"]), box_size_b=s["box_size"].shape[0],
            ray_start=_ptr(s["ray_start"]), ray_dir=_ptr(s["ray_dir"]),
            ray_b=s["ray_start"].shape[0] if s["ray_start"] is not None else 1,
            cam_matrix=_ptr(s["cam_matrix"]), cam_b=s["cam_matrix"].shape[0] if s["cam_matrix"] is not None else 1,
            fov_y=float(scene.get("fov_y", 0.0)),
            step=_ptr(st3), step_b=st3.shape[0], step_h=st3.shape[1], step_w=st3.shape[2],
            tf=_ptr(s["tf"]), tf_b=s["tf"].shape[0], R=s["tf"].shape[1], C=s["tf"].shape[2],
            early_out=float(scene.get("early_out", 1.0 - 1e-5)),
        )
        self.W, self.H = W, H


def forward(scene, dtype=np.float64, frozen_term=None, frozen_t0=None, return_t0=False):
    """Oracle forward render -> color (B,H,W,4), term (B,H,W) int32[, t0 (B,H,W)]."""
    p = _Prepared(scene, dtype)
    B, H, W = p.B, p.H, p.W
    color = np.zeros((B, H, W, 4), dtype=p.dtype)
    term = np.zeros((B, H, W), dtype=np.int32)
    ft = None if frozen_term is None else np.ascontiguousarray(frozen_term, dtype=np.int32)
    f0 = None if frozen_t0 is None else np.ascontiguousarray(frozen_t0, dtype=p.dtype)
    t0 = np.zeros((B, H, W), dtype=p.dtype) if return_t0 else None
    getattr(lib(), f"oracle_render_forward_{_suffix(dtype)}")(
        ctypes.byref(p.s), _ptr(color), _ptr(term), _ptr(ft), _ptr(f0), _ptr(t0))
    return (color, term, t0) if return_t0 else (color, term)


def camera_rays(scene, dtype=np.float64):
    """The scene's camera rays of every pixel (renderer_camera.cuh:12-61) -> start, dir (B,H,W,3)."""
    p = _Prepared(scene, dtype)
    start = np.zeros((p.B, p.H, p.W, 3), dtype=p.dtype)
    d = np.zeros_like(start)
    getattr(lib(), f"oracle_camera_rays_{_suffix(dtype)}")(ctypes.byref(p.s), _ptr(start), _ptr(d))
    return start, d


def forward_gradients(scene, D, d_step=-1, d_start=(-1, -1, -1), d_dir=(-1, -1, -1), d_tf=None, dtype=np.float64,
                      frozen_term=None, frozen_t0=None):
    """Oracle forward-mode derivatives -> color (B,H,W,4), term (B,H,W), grad (B,H,W,D,4).

    d_tf: int array (1 or B, R, C) of derivative indices (-1 = not traced) or None."""
    p = _Prepared(scene, dtype)
    B, H, W = p.B, p.H, p.W
    color = np.zeros((B, H, W, 4), dtype=p.dtype)
    term = np.zeros((B, H, W), dtype=np.int32)
    grad = np.zeros((B, H, W, int(D), 4), dtype=p.dtype)
    ds = np.ascontiguousarray(d_start, dtype=np.int32)
    dd = np.ascontiguousarray(d_dir, dtype=np.int32)
    dt = None if d_tf is None else np.ascontiguousarray(d_tf, dtype=np.int32)
    ft = None if frozen_term is None else np.ascontiguousarray(frozen_term, dtype=np.int32)
    f0 = None if frozen_t0 is None else np.ascontiguousarray(frozen_t0, dtype=p.dtype)
    getattr(lib(), f"oracle_render_forward_gradients_{_suffix(dtype)}")(
        ctypes.byref(p.s), int(D), int(d_step), _ptr(ds), _ptr(dd), _ptr(dt), 0 if dt is None else dt.shape[0],
        _ptr(color), _ptr(term), _ptr(grad), _ptr(ft), _ptr(f0))
    return color, term, grad


def adjoint(scene, color, term, adj_color, grads=("tf",), dtype=np.float64, method=INVERSION, frozen_t0=None,
            adj_volume_batch=None, adj_tf_batch=None, adj_cam_batch=None, adj_step_shape=None):
    """Oracle adjoint -> dict with the requested gradients (accumulated from zero)."""
    p = _Prepared(scene, dtype)
    B, H, W = p.B, p.H, p.W
    vol = p.arrays["volume"]
    tf = p.arrays["tf"]
    out = {}
    s = _AdjOut()
    if "volume" in grads:
        vb = adj_volume_batch if adj_volume_batch is not None else vol.shape[0]
        out["volume"] = np.zeros((vb,) + vol.shape[1:], dtype=p.dtype)
        s.adj_volume, s.adj_volume_b = _ptr(out["volume"]), vb
    if "tf" in grads:
        tb = adj_tf_batch if adj_tf_batch is not None else tf.shape[0]
        out["tf"] = np.zeros((tb,) + tf.shape[1:], dtype=p.dtype)
        s.adj_tf, s.adj_tf_b = _ptr(out["tf"]), tb
    if "camera" in grads:
        cb = adj_cam_batch if adj_cam_batch is not None else B
        out["ray_start"] = np.zeros((cb, H, W, 3), dtype=p.dtype)
        out["ray_dir"] = np.zeros((cb, H, W, 3), dtype=p.dtype)
        s.adj_ray_start, s.adj_ray_dir, s.adj_cam_b = _ptr(out["ray_start"]), _ptr(out["ray_dir"]), cb
    if "step" in grads:
        shp = adj_step_shape if adj_step_shape is not None else (B, H, W)
        out["step"] = np.zeros(shp, dtype=p.dtype)
        s.adj_step, s.adj_step_b, s.adj_step_h, s.adj_step_w = _ptr(out["step"]), shp[0], shp[1], shp[2]
    c = np.ascontiguousarray(color, dtype=p.dtype)
    t = np.ascontiguousarray(term, dtype=np.int32)
    a = np.ascontiguousarray(adj_color, dtype=p.dtype)
    f0 = None if frozen_t0 is None else np.ascontiguousarray(frozen_t0, dtype=p.dtype)
    getattr(lib(), f"oracle_render_adjoint_{_suffix(dtype)}")(
        ctypes.byref(p.s), _ptr(c), _ptr(t), _ptr(a), ctypes.byref(s), int(method), _ptr(f0))
    return out


def kink_census(scene, term, eps=1e-5, eps_pos=2e-5, eps_face=0.0):
    """Per-ray count (B,H,W) of samples near a kink of the reference's mathematics,
    evaluated in FP64 with the given sample counts: within eps * (cell magnitude)
    of a TF clamp / texel edge / control point or of the alpha clamp, within
    eps_pos voxels of a nearest-neighbour rounding edge, or (eps_face > 0) within
    eps_face voxels of an interior trilinear cell face, where the position
    derivative jumps."""
    p = _Prepared(scene, np.float64)
    t = np.ascontiguousarray(term, dtype=np.int32)
    out = np.zeros((p.B, p.H, p.W), dtype=np.int32)
    lib().oracle_kink_census_f64(ctypes.byref(p.s), _ptr(t), ctypes.c_double(eps), ctypes.c_double(eps_pos),
                                 ctypes.c_double(eps_face), _ptr(out))
    return out


# ---------------------------------------------------------------------------
# components
# ---------------------------------------------------------------------------
def _arr(x, dtype):
    return np.ascontiguousarray(np.asarray(x, dtype=dtype))


def blend(mode, acc, cur, step, dtype=np.float64):
    out = np.zeros(4, dtype=dtype)
    getattr(lib(), f"oracle_blend_{_suffix(dtype)}")(int(mode), _ptr(_arr(acc, dtype)), _ptr(_arr(cur, dtype)),
                                                      _c(dtype)(step), _ptr(out))
    return out


def blend_adjoint(mode, acc_out, cur, step, adj_out, dtype=np.float64):
    """Inversion adjoint of one blend step -> (acc_in, adj_acc_in, adj_cur, adj_step)."""
    acc_in, adj_acc, adj_cur = (np.zeros(4, dtype=dtype) for _ in range(3))
    adj_step = np.zeros(1, dtype=dtype)
    getattr(lib(), f"oracle_blend_adjoint_{_suffix(dtype)}")(
        int(mode), _ptr(_arr(acc_out, dtype)), _ptr(_arr(cur, dtype)), _c(dtype)(step), _ptr(_arr(adj_out, dtype)),
        _ptr(acc_in), _ptr(adj_acc), _ptr(adj_cur), _ptr(adj_step))
    return acc_in, adj_acc, adj_cur, float(adj_step[0])


def tf_eval(mode, tf, density, dtype=np.float64):
    tf = _arr(tf, dtype)
    out = np.zeros(4, dtype=dtype)
    getattr(lib(), f"oracle_tf_eval_{_suffix(dtype)}")(int(mode), _ptr(tf), tf.shape[0], tf.shape[1],
                                                        _c(dtype)(density), _ptr(out))
    return out


def tf_adjoint(mode, tf, density, adj, want_tf=True, dtype=np.float64):
    tf = _arr(tf, dtype)
    g = np.zeros_like(tf) if want_tf else None
    d = getattr(lib(), f"oracle_tf_adjoint_{_suffix(dtype)}")(int(mode), _ptr(tf), tf.shape[0], tf.shape[1],
                                                               _c(dtype)(density), _ptr(_arr(adj, dtype)), _ptr(g))
    return float(d), g


def fetch(filter_mode, vol, p, dtype=np.float64):
    vol = _arr(vol, dtype)
    return float(getattr(lib(), f"oracle_fetch_{_suffix(dtype)}")(int(filter_mode), _ptr(vol), *vol.shape,
                                                                   _ptr(_arr(p, dtype))))


def trilinear_adjoint(vol, p, adj_d, want_volume=True, dtype=np.float64):
    vol = _arr(vol, dtype)
    g = np.zeros_like(vol) if want_volume else None
    ap = np.zeros(3, dtype=dtype)
    getattr(lib(), f"oracle_trilinear_adjoint_{_suffix(dtype)}")(_ptr(vol), *vol.shape, _ptr(_arr(p, dtype)),
                                                                 _c(dtype)(adj_d), _ptr(g), _ptr(ap))
    return ap, g


def adj_cross(a, b, adj_c, dtype=np.float64):
    aa, ab = np.zeros(3, dtype=dtype), np.zeros(3, dtype=dtype)
    getattr(lib(), f"oracle_adj_cross_{_suffix(dtype)}")(_ptr(_arr(a, dtype)), _ptr(_arr(b, dtype)),
                                                          _ptr(_arr(adj_c, dtype)), _ptr(aa), _ptr(ab))
    return aa, ab


def adj_normalize(a, adj_out, dtype=np.float64):
    r = np.zeros(3, dtype=dtype)
    getattr(lib(), f"oracle_adj_normalize_{_suffix(dtype)}")(_ptr(_arr(a, dtype)), _ptr(_arr(adj_out, dtype)),
                                                              _ptr(r))
    return r


def generate_rays(viewport, H, W, fov_y, dtype=np.float64):
    vp = _arr(viewport, dtype)
    B = vp.shape[0]
    rs, rd = np.zeros((B, H, W, 3), dtype=dtype), np.zeros((B, H, W, 3), dtype=dtype)
    getattr(lib(), f"oracle_generate_rays_{_suffix(dtype)}")(_ptr(vp), B, H, W, _c(dtype)(fov_y), _ptr(rs), _ptr(rd))
    return rs, rd


def generate_rays_adjoint(viewport, H, W, fov_y, adj_start, adj_dir, dtype=np.float64):
    vp = _arr(viewport, dtype)
    out = np.zeros_like(vp)
    getattr(lib(), f"oracle_generate_rays_adjoint_{_suffix(dtype)}")(
        _ptr(vp), vp.shape[0], H, W, _c(dtype)(fov_y), _ptr(_arr(adj_start, dtype)), _ptr(_arr(adj_dir, dtype)),
        _ptr(out))
    return out


# ---------------------------------------------------------------------------
# auxiliary kernels (renderer/tf_utils_impl.cu, renderer/losses_impl.cu).
# Volumes are dense (X, Y, Z) / (4, X, Y, Z) arrays, tf an (R, C) table.
# ---------------------------------------------------------------------------
def preshade(mode, tf, density, dtype=np.float64):
    """colour (4, X, Y, Z) = TF(density) per voxel (renderer/tf_utils_impl.cu:17-32)."""
    tf, d = _arr(tf, dtype), _arr(density, dtype)
    out = np.zeros((4,) + d.shape, dtype=dtype)
    getattr(lib(), f"oracle_preshade_{_suffix(dtype)}")(int(mode), _ptr(tf), tf.shape[0], tf.shape[1], _ptr(d),
                                                         ctypes.c_int64(d.size), _ptr(out))
    return out


def preshade_adjoint(mode, tf, density, adj_color, dtype=np.float64):
    """d<adj_color, TF(density)>/d density per voxel (renderer/tf_utils_impl.cu:96-121)."""
    tf, d, a = _arr(tf, dtype), _arr(density, dtype), _arr(adj_color, dtype)
    out = np.zeros(d.shape, dtype=dtype)
    getattr(lib(), f"oracle_preshade_adjoint_{_suffix(dtype)}")(int(mode), _ptr(tf), tf.shape[0], tf.shape[1],
                                                                 _ptr(d), ctypes.c_int64(d.size), _ptr(a), _ptr(out))
    return out


def splitmix64(x: int) -> int:
    return int(lib().oracle_splitmix64(ctypes.c_uint64(x & (2 ** 64 - 1))))


def lcg48_next(state: int) -> int:
    return int(lib().oracle_lcg48_next(ctypes.c_uint64(state)))


def best_fit_candidates(seed: int, voxel: int, n: int) -> np.ndarray:
    """The n candidate densities of voxel index (x*Y + y)*Z + z."""
    s = int(lib().oracle_best_fit_stream(ctypes.c_uint64(seed), ctypes.c_uint64(voxel)))
    out = np.empty(n)
    for k in range(n):
        s = lcg48_next(s)
        out[k] = (s >> 24) / 16777216.0
    return out


def best_fit(mode, tf, color, num_samples, opacity_weighting, prev=None, neighbor_weighting=0.0, seed=0,
             dtype=np.float64):
    """Best-fit densities (X, Y, Z) and their costs (renderer/tf_utils_impl.cu:225-272, :351-420)."""
    tf, c = _arr(tf, dtype), _arr(color, dtype)
    pv = None if prev is None else _arr(prev, dtype)
    X, Y, Z = c.shape[1:]
    out, cost = np.zeros((X, Y, Z), dtype=dtype), np.zeros((X, Y, Z), dtype=dtype)
    ct = _c(dtype)
    getattr(lib(), f"oracle_best_fit_{_suffix(dtype)}")(
        int(mode), _ptr(tf), tf.shape[0], tf.shape[1], _ptr(c), X, Y, Z, int(num_samples), ct(opacity_weighting),
        _ptr(pv), ct(neighbor_weighting), ctypes.c_uint64(seed), _ptr(out), _ptr(cost))
    return out, cost


def best_fit_cost(mode, tf, target, d, opacity_weighting, nb=None, has=None, neighbor_weighting=0.0,
                  dtype=np.float64):
    tf, t = _arr(tf, dtype), _arr(target, dtype)
    nbv = None if nb is None else _arr(nb, dtype)
    hv = None if has is None else np.ascontiguousarray(np.asarray(has, dtype=np.int32))
    ct = _c(dtype)
    return float(getattr(lib(), f"oracle_best_fit_cost_{_suffix(dtype)}")(
        int(mode), _ptr(tf), tf.shape[0], tf.shape[1], _ptr(t), ct(d), ct(opacity_weighting), _ptr(nbv), _ptr(hv),
        ct(neighbor_weighting)))


def importance_thresholds(weights):
    """renderer/losses.cpp:188-200 in float32: cumulative weights divided by their sum."""
    u, dg, o, og = (np.float32(w) for w in weights)
    dg = np.float32(dg + u)
    o = np.float32(o + dg)
    og = np.float32(og + o)
    return np.array([u / og, dg / og, o / og], dtype=np.float32)


def importance_maxima(mode, tf, density, dtype=np.float64):
    """max |grad density|, max opacity, max |grad opacity| (renderer/losses_impl.cu:254-292)."""
    tf, d = _arr(tf, dtype), _arr(density, dtype)
    out = np.zeros(3, dtype=dtype)
    X, Y, Z = d.shape
    getattr(lib(), f"oracle_importance_maxima_{_suffix(dtype)}")(int(mode), _ptr(tf), tf.shape[0], tf.shape[1],
                                                                  _ptr(d), X, Y, Z, _ptr(out))
    return out


def sample_importance(mode, tf, density, locations, weights, maxima, seed, dtype=np.float64):
    """Decisions (N,) and their margins |value/max - rndReject| (inf for the uniform
    branch) with the GPU kernel's random numbers (renderer/losses_impl.cu:298-371)."""
    tf, d, loc = _arr(tf, dtype), _arr(density, dtype), _arr(locations, dtype)
    mx = _arr(maxima, dtype)
    w = np.ascontiguousarray(importance_thresholds(weights))
    N = loc.shape[0]
    out = np.zeros(N, dtype=np.uint8)
    margin = np.zeros(N, dtype=dtype)
    X, Y, Z = d.shape
    getattr(lib(), f"oracle_sample_importance_{_suffix(dtype)}")(
        int(mode), _ptr(tf), tf.shape[0], tf.shape[1], _ptr(d), X, Y, Z, _ptr(loc), ctypes.c_int64(N), _ptr(w),
        _ptr(mx), ctypes.c_uint32(seed & 0xFFFFFFFF), _ptr(out), _ptr(margin))
    return out.astype(bool), margin


# elementwise losses (renderer/losses_impl.cu:35-240 with renderer/renderer_utils.cuh:180-298), FP64
_FLT_EPSILON = 1.192092896e-07


def _log1mexp(u):
    """log(1 - e^u) for u <= 0 without cancellation (Maechler 2012)."""
    u = np.asarray(u, dtype=np.float64)
    with np.errstate(divide="ignore"):
        return np.where(u > -np.log(2.0), np.log(-np.expm1(np.minimum(u, 0.0))),
                        np.log1p(-np.exp(np.minimum(u, 0.0))))


def mul_log(t):
    t = np.asarray(t, dtype=np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(t < _FLT_EPSILON, 0.0, t * np.log2(np.maximum(t, _FLT_EPSILON)))


def mul_log_backward(t, grad):
    t = np.asarray(t, dtype=np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(t < _FLT_EPSILON, 0.0, grad * (np.log2(np.maximum(t, _FLT_EPSILON)) + 1.0 / np.log(2.0)))


def log_mse(log_x, log_y):
    a, b = np.maximum(log_x, log_y), np.minimum(log_x, log_y)
    return 2.0 * (_log1mexp(b - a) + a)


def log_l1(log_x, log_y):
    a, b = np.maximum(log_x, log_y), np.minimum(log_x, log_y)
    return _log1mexp(b - a) + a


def log_mse_backward(log_x, log_y, grad):
    with np.errstate(divide="ignore"):
        return 2.0 * grad / -np.expm1(log_y - log_x), 2.0 * grad / -np.expm1(log_x - log_y)


def log_l1_backward(log_x, log_y, grad):
    with np.errstate(divide="ignore"):
        return grad / -np.expm1(log_y - log_x), grad / -np.expm1(log_x - log_y)


# ---------------------------------------------------------------------------
# camera viewports (renderer/camera.cpp:127-198, orientation tables :70-85),
# numpy float64; complex inputs are carried through unchanged so that the
# complex-step derivative (Im f(x + ih) / h) gives exact Jacobians
# ---------------------------------------------------------------------------
ORIENTATION_UP = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], dtype=np.float64)
ORIENTATION_PERMUTATION = [(2, -1, -3), (-2, 1, 3), (1, 2, 3), (-1, -2, -3), (-3, -1, 2), (3, 1, -2)]
ORIENTATION_INVERT_YAW = [True, False, False, True, False, True]
ORIENTATION_INVERT_PITCH = [False, False, False, False, False, False]


def _cross(a, b):
    return np.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1],
                     a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                     a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], axis=-1)


def _normalize(v):
    # sqrt of the sum of squares (analytic: no abs), as torch.nn.functional.normalize away from 0
    return v / np.sqrt(np.sum(v * v, axis=-1, keepdims=True))


def viewport_from_lookat(origin, look_at, up):
    """B x 3 x 3 rows (eye, right, up) (camera.cpp:127-147): front = normalize(lookAt - origin),
    right = normalize(front x up), up' = normalize(right x front)."""
    origin, look_at, up = (np.atleast_2d(np.asarray(a)) for a in (origin, look_at, up))
    B = max(origin.shape[0], look_at.shape[0], up.shape[0])
    origin, look_at, up = (np.broadcast_to(a, (B, 3)) for a in (origin, look_at, up))
    front = _normalize(look_at - origin)
    right = _normalize(_cross(front, up))
    up2 = _normalize(_cross(right, front))
    return np.stack([origin, right, up2], axis=1)


def viewport_from_sphere(center, yaw, pitch, distance, orientation):
    """Viewport of a camera at `distance` from `center` (camera.cpp:149-198): yaw and pitch are
    negated unless the orientation inverts them; position (cos p cos y, sin p, cos p sin y) * r is
    permuted / sign-flipped by the orientation's permutation; looks at the center with the
    orientation's up vector.  center B x 3, yaw / pitch / distance B x 1 (size-1 batches broadcast)."""
    o = int(orientation)
    center = np.atleast_2d(np.asarray(center))
    yaw = np.asarray(yaw).reshape(-1, 1) * (1.0 if ORIENTATION_INVERT_YAW[o] else -1.0)
    pitch = np.asarray(pitch).reshape(-1, 1) * (1.0 if ORIENTATION_INVERT_PITCH[o] else -1.0)
    distance = np.asarray(distance).reshape(-1, 1)
    pos = [np.cos(pitch) * np.cos(yaw) * distance, np.sin(pitch) * distance, np.cos(pitch) * np.sin(yaw) * distance]
    pos2 = [pos[abs(p) - 1] * (1.0 if p > 0 else -1.0) for p in ORIENTATION_PERMUTATION[o]]
    B = max(center.shape[0], yaw.shape[0], pitch.shape[0], distance.shape[0])
    origin = np.broadcast_to(np.concatenate(np.broadcast_arrays(*pos2), axis=1), (B, 3))
    c = np.broadcast_to(center, (B, 3))
    return viewport_from_lookat(origin + c, c, np.broadcast_to(ORIENTATION_UP[o], (B, 3)))


def viewport_from_sphere_jacobian(center, yaw, pitch, distance, orientation, h=1e-30):
    """d viewport / d (yaw, pitch, distance): B x 3 x 3 x 3 (last axis: yaw, pitch, distance),
    by the complex step (exact to rounding for these analytic functions)."""
    args = [np.asarray(yaw, dtype=np.float64).reshape(-1, 1), np.asarray(pitch, dtype=np.float64).reshape(-1, 1),
            np.asarray(distance, dtype=np.float64).reshape(-1, 1)]
    cols = []
    for k in range(3):
        a = [x.astype(np.complex128) for x in args]
        a[k] = a[k] + 1j * h
        cols.append(np.imag(viewport_from_sphere(np.asarray(center, dtype=np.float64), a[0], a[1], a[2],
                                                 orientation)) / h)
    return np.stack(cols, axis=-1)
