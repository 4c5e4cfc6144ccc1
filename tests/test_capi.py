"""The C-ABI boundary on the CPU (no GPU calls): libdvr_hip.so loads, exports
every function include/dvr.h declares, and rejects invalid calls with a
status code and a message (renderer/renderer.cpp:48-391 validation order),
never silently (the reference ignores launch errors, renderer_cuda.cpp:109-110).
Every call here fails validation before any HIP call is made.
"""
import ctypes
import os
import re

import pytest
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dvr.h")
LIB = os.path.join(ROOT, "diffdvr_amd", "libdvr_hip.so")

# these tests fabricate device descriptors; never let them near a real device
pytestmark = pytest.mark.skipif(torch.cuda.is_available(), reason="validation-only tests run where no GPU is visible")


class Tensor(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("sizes", ctypes.c_int64 * 6), ("strides", ctypes.c_int64 * 6),
                ("ndim", ctypes.c_int32), ("dtype", ctypes.c_int32), ("device", ctypes.c_int32)]


class Inputs(ctypes.Structure):
    _fields_ = [("screen_width", ctypes.c_int32), ("screen_height", ctypes.c_int32),
                ("volume_filter_mode", ctypes.c_int32), ("volume", Tensor),
                ("box_min", Tensor), ("box_min_value", ctypes.c_float * 3),
                ("box_size", Tensor), ("box_size_value", ctypes.c_float * 3),
                ("camera_mode", ctypes.c_int32), ("camera_ray_start", Tensor), ("camera_ray_dir", Tensor),
                ("camera_matrix", Tensor), ("camera_matrix_value", ctypes.c_float * 16),
                ("camera_fov_y_radians", ctypes.c_float), ("step_size", Tensor), ("step_size_value", ctypes.c_float),
                ("tf_mode", ctypes.c_int32), ("tf", Tensor), ("blend_mode", ctypes.c_int32),
                ("blending_early_out", ctypes.c_float)]


class Outputs(ctypes.Structure):
    _fields_ = [("color", Tensor), ("termination_index", Tensor)]


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:dvr_status|int32_t|int64_t|const char\s*\*)\s*(dvr_\w+)\s*\(", text,
                                 flags=re.M)))


def lib():
    l = ctypes.CDLL(LIB)
    l.dvr_last_error.restype = ctypes.c_char_p
    l.dvr_version.restype = ctypes.c_char_p
    l.dvr_checkpoint_buffer_elements.restype = ctypes.c_int64
    return l


def tensor(shape, device, dtype=0, data=0x1000):
    t = Tensor()
    t.data = data
    t.ndim = len(shape)
    t.dtype = dtype
    t.device = device
    stride = 1
    for d in reversed(range(len(shape))):
        t.sizes[d] = shape[d]
        t.strides[d] = stride
        stride *= shape[d]
    return t


def valid_inputs(device):
    i = Inputs()
    i.screen_width, i.screen_height = 8, 6
    i.volume_filter_mode = 1
    i.volume = tensor((1, 5, 5, 5), device)
    i.box_min_value[:] = [-0.5, -0.5, -0.5]
    i.box_size_value[:] = [1, 1, 1]
    i.camera_mode = 2
    i.camera_matrix = tensor((1, 3, 3), device)
    i.camera_fov_y_radians = 0.7
    i.step_size_value = 0.01
    i.tf_mode = 1
    i.tf = tensor((1, 4, 4), device)
    i.blend_mode = 0
    i.blending_early_out = 1 - 1e-5
    return i


def test_header_functions_are_exported():
    names = declared_functions()
    assert len(names) >= 14, names
    l = lib()
    missing = [n for n in names if not hasattr(l, n)]
    assert not missing, missing


def test_library_identity():
    l = lib()
    assert l.dvr_use_double_precision() == 0
    assert b"gfx950" in l.dvr_version()


def test_host_tensors_are_rejected_loudly():
    l = lib()
    i = valid_inputs(-1)
    o = Outputs(tensor((1, 6, 8, 4), -1), tensor((1, 6, 8), -1, dtype=1))
    st = l.dvr_render_forward(ctypes.byref(i), ctypes.byref(o), None)
    assert st == 3  # DVR_ERR_UNSUPPORTED
    assert b"GPU only" in l.dvr_last_error()


@pytest.mark.parametrize("mutate,code,msg", [
    (lambda i: setattr(i, "tf", tensor((1, 4, 5), 0)), 1, b"tf must have 4 entries at dimension 2"),
    (lambda i: setattr(i, "volume", tensor((1, 5, 5, 5), 0, dtype=1)), 1, b"volume must be of type torch.float32"),
    (lambda i: setattr(i, "camera_matrix", tensor((2, 3, 3), 0)) or setattr(i, "tf", tensor((3, 4, 4), 0)), 1,
     b"must be have the same batch size"),
    (lambda i: setattr(i, "camera_mode", 7), 1, b"unknown camera enum value"),
    (lambda i: setattr(i, "tf_mode", 6), 1, b"unknown tf enum value"),
    (lambda i: setattr(i, "tf_mode", 5), 1, b"both the filter mode and TF mode must be set to preshaded"),
    (lambda i: setattr(i, "volume_filter_mode", 3) or setattr(i, "tf_mode", 5), 1,
     b"volume colors are encoded in the batch dimension of the volume. It must be 4"),
    (lambda i: setattr(i, "step_size_value", 0.0), 1, b"step size must be positive"),
    (lambda i: setattr(i, "volume", tensor((1, 1, 5, 5), 0)), 1, b"at least 2 voxels"),
])
def test_validation_messages(mutate, code, msg):
    l = lib()
    i = valid_inputs(0)
    mutate(i)
    o = Outputs(tensor((1, 6, 8, 4), 0), tensor((1, 6, 8), 0, dtype=1))
    st = l.dvr_render_forward(ctypes.byref(i), ctypes.byref(o), None)
    assert st == code, (st, l.dvr_last_error())
    assert msg in l.dvr_last_error()


class FwdOptions(ctypes.Structure):
    _fields_ = [("path", ctypes.c_int32), ("checkpoint_interval", ctypes.c_int32), ("checkpoints", Tensor),
                ("densities", Tensor)]


@pytest.mark.parametrize("path,K", [(2, 0), (0, 16)])
def test_preshaded_renders_on_the_plain_path_only(path, K):
    """A pre-shaded volume (4 colour channels) with the tiled path or checkpoints asked for is
    refused before any HIP call, naming the plain path."""
    l = lib()
    i = valid_inputs(0)
    i.volume_filter_mode, i.tf_mode = 3, 5
    i.volume = tensor((4, 5, 5, 5), 0)
    o = Outputs(tensor((1, 6, 8, 4), 0), tensor((1, 6, 8), 0, dtype=1))
    opt = FwdOptions(path, K, tensor((1 << 20,), 0))
    assert l.dvr_render_forward_ex(ctypes.byref(i), ctypes.byref(o), ctypes.byref(opt), None) == 3
    assert b"pre-shaded volumes render on the plain path only" in l.dvr_last_error()


def test_output_shape_mismatch_and_null_arguments():
    l = lib()
    i = valid_inputs(0)
    o = Outputs(tensor((1, 6, 7, 4), 0), tensor((1, 6, 8), 0, dtype=1))
    assert l.dvr_render_forward(ctypes.byref(i), ctypes.byref(o), None) == 1
    assert b"color must have 8 entries at dimension 2" in l.dvr_last_error()
    assert l.dvr_render_forward(None, ctypes.byref(o), None) == 1
    assert l.dvr_checkpoint_buffer_elements(None, 16) == -1


# ---------------------------------------------------------------------------
# batched plans: validation happens before any HIP call
# ---------------------------------------------------------------------------
def _plan(volumes=None, tfs=None, views=None, tf_mode=1, K=16, ww=8, loss=0, forward_only=False, device=0):
    from diffdvr_amd._native import native as n
    volumes = volumes if volumes is not None else [(0x1000, 5, 5, 5, [-0.5] * 3, [1.0] * 3, 0x2000)]
    tfs = tfs if tfs is not None else [(0x3000, 4, 4, 0x4000)]
    views = views if views is not None else [(0, 0, 8, 6, 2, 0x5000, 0.7, 0.01, 0x6000, 0x7000, 0, 0x8000, 0.0, 0)]
    return n._BatchPlan(volumes, tfs, views, 1, tf_mode, 0, 1 - 1e-5, loss, 0, forward_only, device, K, ww, 0)


@pytest.mark.parametrize("kw,msg", [
    (dict(tfs=[(0x3000, 4, 5, 0)]), "tf 0 must have 4 entries at dimension 2"),
    (dict(views=[(3, 0, 8, 6, 2, 0x5000, 0.7, 0.01, 0x6000, 0x7000, 0, 0x8000, 0.0, 0)]), "volume index out of range"),
    (dict(views=[(0, 0, 8, 6, 0, 0x5000, 0.7, 0.01, 0x6000, 0x7000, 0, 0x8000, 0.0, 0)]),
     "ReferenceFrame or InverseViewMatrix"),
    (dict(views=[(0, 0, 8, 6, 2, 0x5000, 0.7, -1.0, 0x6000, 0x7000, 0, 0x8000, 0.0, 0)]), "step size must be positive"),
    (dict(views=[(0, 0, 8, 6, 2, 0x5000, 0.7, 0.01, 0x6000, 0x7000, 0, 0, 0.0, 0)]), "adj_color is NULL"),
    (dict(loss=1), "target is NULL"),
    (dict(volumes=[(0x1000, 1, 5, 5, [-0.5] * 3, [1.0] * 3, 0)]), "at least 2 voxels"),
    (dict(K=12), "checkpoint_interval must be 4, 8, 16, 32 or 64"),
    (dict(ww=6), "wave_width must be 4, 8, 16 or 32"),
    (dict(tf_mode=6), "tf mode 6 is not supported"),
    (dict(device=-1), "GPU only"),
    (dict(views=[(0, 0, 8, 6, 1, 0x5000, 0.7, 0.01, 0x6000, 0x7000, 0, 0x8000, 0.0, 0x9000)]),
     "camera gradients need a ReferenceFrame camera"),
])
def test_batch_plan_validation(kw, msg):
    with pytest.raises(RuntimeError, match=msg):
        _plan(**kw)


def test_adam_group_validation():
    from diffdvr_amd._native import native as n
    with pytest.raises(RuntimeError, match="NULL buffer"):
        n._AdamGroup([(0x1000, 0, 0x2000, 0x3000, 1, 4, 4, 1.0, 0.0)], 0)
    with pytest.raises(RuntimeError, match="invalid shape"):
        n._AdamGroup([(0x1000, 0x1100, 0x2000, 0x3000, 2, 5, 4, 1.0, 0.0)], 0)


# ---------------------------------------------------------------------------
# forward-mode differentiation: validation happens before any HIP call
# ---------------------------------------------------------------------------
class FwdDiff(ctypes.Structure):
    _fields_ = [("D", ctypes.c_int32), ("d_stepsize", ctypes.c_int32), ("d_ray_start", ctypes.c_int32 * 3),
                ("d_ray_dir", ctypes.c_int32 * 3), ("has_tf_derivatives", ctypes.c_int32), ("d_tf", Tensor)]


class AdjOut(ctypes.Structure):
    _fields_ = [("has_volume_derivatives", ctypes.c_int32), ("adj_volume", Tensor),
                ("has_stepsize_derivatives", ctypes.c_int32), ("adj_stepsize", Tensor),
                ("has_camera_derivatives", ctypes.c_int32), ("adj_camera_ray_start", Tensor),
                ("adj_camera_ray_dir", Tensor), ("has_tf_derivatives", ctypes.c_int32), ("adj_tf", Tensor),
                ("tf_delayed_accumulation", ctypes.c_int32)]


def fwd_diff(D=1, step=0, tf=False):
    f = FwdDiff()
    f.D, f.d_stepsize = D, step
    f.d_ray_start[:] = [-1, -1, -1]
    f.d_ray_dir[:] = [-1, -1, -1]
    if tf:
        f.has_tf_derivatives = 1
        f.d_tf = tensor((1, 4, 4), 0, dtype=1)
    return f


@pytest.mark.parametrize("mutate,code,msg", [
    (lambda f, g: setattr(f, "D", 0), 1, b"number of derivatives must be positive"),
    (lambda f, g: setattr(f, "d_stepsize", -1), 1, b"at least one derivative must be activated"),
    (lambda f, g: setattr(f, "d_stepsize", 3), 1, b"d_stepsize = 3 is not a derivative index"),
    (lambda f, g: f.d_ray_dir.__setitem__(1, 1), 1, b"d_rayDir.y = 1 is not a derivative index"),
    (lambda f, g: (setattr(f, "has_tf_derivatives", 1), setattr(f, "d_tf", tensor((1, 4, 4), 0, dtype=0))), 1,
     b"d_tf must be of type torch.int32"),
    (lambda f, g: (setattr(f, "has_tf_derivatives", 1), setattr(f, "d_tf", tensor((1, 4, 3), 0, dtype=1))), 1,
     b"d_tf must have 4 entries at dimension 2"),
    (lambda f, g: setattr(g, "sizes", (ctypes.c_int64 * 6)(1, 6, 8, 2, 4, 0)), 1,
     b"gradients_out must have 1 entries at dimension 3"),
])
def test_forward_gradients_validation(mutate, code, msg):
    l = lib()
    i = valid_inputs(0)
    o = Outputs(tensor((1, 6, 8, 4), 0), tensor((1, 6, 8), 0, dtype=1))
    f = fwd_diff()
    g = tensor((1, 6, 8, 1, 4), 0)
    mutate(f, g)
    st = l.dvr_render_forward_gradients(ctypes.byref(i), ctypes.byref(f), ctypes.byref(o), ctypes.byref(g), None)
    assert st == code, (st, l.dvr_last_error())
    assert msg in l.dvr_last_error()


def test_forward_variables_to_gradients_validation():
    l = lib()
    fv = tensor((1, 6, 8, 2, 4), 0)
    ac = tensor((1, 6, 8, 4), 0)
    a = AdjOut()
    a.has_volume_derivatives = 1
    a.adj_volume = tensor((1, 5, 5, 5), 0)
    f = fwd_diff(D=2)
    assert l.dvr_forward_variables_to_gradients(ctypes.byref(fv), ctypes.byref(ac), ctypes.byref(f), ctypes.byref(a),
                                                None) == 1
    assert b"volume derivatives are not supported" in l.dvr_last_error()
    a = AdjOut()
    f = fwd_diff(D=3)
    assert l.dvr_forward_variables_to_gradients(ctypes.byref(fv), ctypes.byref(ac), ctypes.byref(f), ctypes.byref(a),
                                                None) == 1
    assert b"forwardVariables has 2 derivatives" in l.dvr_last_error()
    f = fwd_diff(D=2)
    assert l.dvr_forward_variables_to_gradients(ctypes.byref(fv), ctypes.byref(ac), ctypes.byref(f), ctypes.byref(a),
                                                None) == 2
    assert b"At least one adjoint variable" in l.dvr_last_error()
    ac_bad = tensor((1, 6, 7, 4), 0)
    assert l.dvr_forward_variables_to_gradients(ctypes.byref(fv), ctypes.byref(ac_bad), ctypes.byref(f),
                                                ctypes.byref(a), None) == 1
    assert b"gradientOutputColor must have 8 entries at dimension 2" in l.dvr_last_error()


def test_compare_to_image_validation():
    l = lib()
    col = tensor((1, 6, 8, 4), 0)
    g = tensor((1, 6, 8, 3, 4), 0)
    ref = tensor((1, 6, 8, 4), 0)
    d1, g1 = tensor((1,), 0), tensor((2,), 0)
    assert l.dvr_compare_to_image(ctypes.byref(col), ctypes.byref(g), ctypes.byref(ref), ctypes.byref(d1),
                                  ctypes.byref(g1), 1, None) == 1
    assert b"gradientsOut must have 3 entries at dimension 0" in l.dvr_last_error()
    d4, g4 = tensor((1, 6, 8, 1), 0), tensor((1, 6, 8, 3), 0)
    ref_bad = tensor((2, 6, 8, 4), 0)
    assert l.dvr_compare_to_image(ctypes.byref(col), ctypes.byref(g), ctypes.byref(ref_bad), ctypes.byref(d4),
                                  ctypes.byref(g4), 0, None) == 1
    assert b"colorReference must have 1 entries at dimension 0" in l.dvr_last_error()
    host = tensor((1, 6, 8, 4), -1)
    assert l.dvr_compare_to_image(ctypes.byref(host), ctypes.byref(g), ctypes.byref(ref), ctypes.byref(d4),
                                  ctypes.byref(g4), 0, None) == 3


# ---------------------------------------------------------------------------
# auxiliary kernels: validation before any HIP call (renderer/tf_utils.cpp:173-334,
# renderer/losses.cpp:8-221)
# ---------------------------------------------------------------------------
def test_preshade_validation_messages():
    l = lib()
    d, c = tensor((1, 4, 4, 4), 0), tensor((4, 4, 4, 4), 0)
    st = l.dvr_preshade_volume(ctypes.byref(d), ctypes.byref(tensor((1, 2, 6), 0)), 5, ctypes.byref(c), None)
    assert st == 1 and b"unknown tf enum value: TFPreshaded" in l.dvr_last_error()
    st = l.dvr_preshade_volume(ctypes.byref(d), ctypes.byref(tensor((1, 3, 4), 0)), 0, ctypes.byref(c), None)
    assert st == 1 and b"tf must have 1 entries at dimension 1" in l.dvr_last_error()
    st = l.dvr_preshade_volume(ctypes.byref(d), ctypes.byref(tensor((1, 128, 256), 0)), 1, ctypes.byref(c), None)
    assert st == 1 and b"tf must have 4 entries at dimension 2" in l.dvr_last_error()
    big = tensor((1, 8192, 6), 0)
    st = l.dvr_preshade_volume(ctypes.byref(d), ctypes.byref(big), 6, ctypes.byref(c), None)
    assert st == 3 and b"at most 16384 entries" in l.dvr_last_error()
    st = l.dvr_preshade_volume(ctypes.byref(tensor((1, 4, 4, 4), -1)), ctypes.byref(tensor((1, 2, 6), -1)), 3,
                               ctypes.byref(tensor((4, 4, 4, 4), -1)), None)
    assert st == 3 and b"GPU only" in l.dvr_last_error()
    st = l.dvr_preshade_volume(ctypes.byref(d), ctypes.byref(tensor((1, 2, 6), 0)), 3,
                               ctypes.byref(tensor((4, 4, 4, 5), 0)), None)
    assert st == 1 and b"color must have 4 entries at dimension 3" in l.dvr_last_error()


def test_best_fit_and_importance_validation_messages():
    l = lib()
    c, tf, out = tensor((4, 5, 5, 5), 0), tensor((1, 8, 4), 0), tensor((1, 5, 5, 5), 0)
    seed = ctypes.c_uint64(1)
    st = l.dvr_find_best_fit(ctypes.byref(c), ctypes.byref(tf), 1, 0, ctypes.c_float(1.0), None,
                             ctypes.c_float(0.0), seed, ctypes.byref(out), None)
    assert st == 1 and b"numSamples must be >= 1, but is 0" in l.dvr_last_error()
    st = l.dvr_find_best_fit(ctypes.byref(c), ctypes.byref(tf), 1, 8, ctypes.c_float(-2.0), None,
                             ctypes.c_float(0.0), seed, ctypes.byref(out), None)
    assert st == 1 and b"opacity weighting must be >= 0" in l.dvr_last_error()
    st = l.dvr_find_best_fit(ctypes.byref(tensor((3, 5, 5, 5), 0)), ctypes.byref(tf), 1, 8, ctypes.c_float(1.0),
                             None, ctypes.c_float(0.0), seed, ctypes.byref(out), None)
    assert st == 1 and b"colorVolume must have 4 entries at dimension 0" in l.dvr_last_error()
    w = (ctypes.c_float * 4)(1, 0, 0, 0)
    mx, flags = tensor((3,), 0), tensor((10,), 0, dtype=3)
    st = l.dvr_sample_importance(ctypes.byref(tensor((10, 3), -1)), ctypes.byref(tensor((1, 5, 5, 5), -1)),
                                 ctypes.byref(tensor((1, 8, 4), -1)), 1, w, 0, ctypes.byref(tensor((3,), -1)),
                                 ctypes.byref(tensor((10,), -1, dtype=3)), None)
    assert st == 3 and b"sampleImportance is only supported on CUDA" in l.dvr_last_error()
    st = l.dvr_sample_importance(ctypes.byref(tensor((10, 3), 0)), ctypes.byref(tensor((1, 5, 5, 5), 0)),
                                 ctypes.byref(tf), 6, w, 0, ctypes.byref(mx), ctypes.byref(flags), None)
    assert st == 1 and l.dvr_last_error() == b"unknown tf enum value"
    wz = (ctypes.c_float * 4)(0, 0, 0, 0)
    st = l.dvr_sample_importance(ctypes.byref(tensor((10, 3), 0)), ctypes.byref(tensor((1, 5, 5, 5), 0)),
                                 ctypes.byref(tf), 1, wz, 0, ctypes.byref(mx), ctypes.byref(flags), None)
    assert st == 1 and b"at least one weight parameter must be positive" in l.dvr_last_error()


def test_elementwise_loss_validation_messages():
    l = lib()
    a, b = tensor((10, 3), 0), tensor((10, 4), 0)
    st = l.dvr_log_mse(ctypes.byref(a), ctypes.byref(b), ctypes.byref(tensor((10, 3), 0)), None)
    assert st == 1 and b"same shape" in l.dvr_last_error()
    st = l.dvr_log_l1(ctypes.byref(a), ctypes.byref(tensor((10, 3), 0, dtype=2)), ctypes.byref(a), None)
    assert st == 1 and b"same dtype" in l.dvr_last_error()
    nc = tensor((10, 3), 0)
    nc.strides[0], nc.strides[1] = 1, 10
    st = l.dvr_mul_log(ctypes.byref(nc), ctypes.byref(tensor((10, 3), 0)), None)
    assert st == 1 and b"t must be contiguous" in l.dvr_last_error()
    st = l.dvr_mul_log(ctypes.byref(tensor((10,), 0, dtype=1)), ctypes.byref(tensor((10,), 0, dtype=1)), None)
    assert st == 1 and b"torch.float32 or torch.float64" in l.dvr_last_error()
