// C-ABI of libdvr_hip.so: validation, parameter packing and kernel dispatch.
//
// Validation follows renderer::Renderer::checkInputs / checkOutput /
// checkAdjointOutput (renderer/renderer.cpp:48-391) and checkBatch
// (renderer/pytorch_utils.h:12-20), with the same messages where the
// reference has them.  Unlike the reference (renderer/renderer_cuda.cpp:109-110,
// renderer/renderer.cpp:407-409) launch failures are reported, never ignored.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <optional>
#include <string>
#include <vector>

#include "../../include/dvr.h"
#include "dvr_launch.hpp"
#include "dvr_params.hpp"
#include "dvr_host_util.hpp"

std::string& dvr_host::last_error()
{
    static thread_local std::string e;
    return e;
}

namespace {

using dvr_host::Fail;
using dvr_host::fail;
using dvr_host::check_hip;
using dvr_host::guarded;
using dvr_host::DeviceGuard;

using dvr_host::check_device;
using dvr_host::check_dim;
using dvr_host::check_dtype;
using dvr_host::check_size;
using dvr_host::given;

// checkBatch: size 1 broadcasts
int check_batch(const dvr_tensor& t, const char* name, int B1, int idx = 0)
{
    const int B2 = (int)t.sizes[idx];
    if (B2 == 1) return B1;
    if (B1 == 1) return B2;
    if (B1 != B2)
        fail(DVR_ERR_INVALID_ARGUMENT,
             "%s must be have the same batch size as other tensors, if batch>1. actual=%d, expected=%d", name, B2, B1);
    return B1;
}

// broadcasting stride: a size-1 dimension gets stride 0 (renderer_tensor.cuh:140-144)
int bstride(const dvr_tensor& t, int d)
{
    if (t.sizes[d] == 1) return 0;
    return (int)t.strides[d];
}

// every element offset must fit into int32
void check_offsets(const dvr_tensor& t, const char* name)
{
    int64_t maxo = 0;
    for (int d = 0; d < t.ndim; ++d) {
        if (t.sizes[d] <= 0) continue;
        if (t.strides[d] < 0) fail(DVR_ERR_INVALID_ARGUMENT, "%s has a negative stride (unsupported)", name);
        maxo += (t.sizes[d] - 1) * t.strides[d];
    }
    if (maxo >= (int64_t(1) << 31)) fail(DVR_ERR_INVALID_ARGUMENT, "%s is too large for 32-bit element offsets", name);
}

// One batch slice of a volume (dims 1..3) is addressed with 32-bit byte
// offsets by the kernels: its largest element offset must stay below 2^30.
void check_slice_offsets(const dvr_tensor& t, const char* name)
{
    int64_t maxo = 0;
    for (int d = 1; d < t.ndim; ++d)
        if (t.sizes[d] > 0) maxo += (t.sizes[d] - 1) * t.strides[d];
    if (maxo >= (int64_t(1) << 30)) fail(DVR_ERR_UNSUPPORTED, "%s: one volume may hold at most 2^30 voxels", name);
}

struct Checked {
    dvr::RenderParams P;
    int B, W, H, X, Y, Z;
    int device;
    int filter, camera, tfm, blend;
    bool preshaded;  // volume holds 4 colour channels in its batch dimension
};

Checked check_inputs(const dvr_inputs* in)
{
    if (!in) fail(DVR_ERR_INVALID_ARGUMENT, "inputs is NULL");
    Checked c{};
    dvr::RenderParams& P = c.P;
    std::memset(&P, 0, sizeof(P));
    c.W = in->screen_width;
    c.H = in->screen_height;
    if (c.W < 0 || c.H < 0) fail(DVR_ERR_INVALID_ARGUMENT, "screen size must be non-negative");

    const dvr_tensor& vol = in->volume;
    check_dim(vol, "volume", 4);
    c.device = vol.device;
    if (c.device < 0)
        fail(DVR_ERR_UNSUPPORTED,
             "volume is a host (CPU) tensor: this library renders on the GPU only; move the inputs to the HIP device");
    check_dtype(vol, "volume", DVR_FLOAT32);
    if (in->volume_filter_mode != DVR_FILTER_NEAREST && in->volume_filter_mode != DVR_FILTER_TRILINEAR &&
        in->volume_filter_mode != DVR_FILTER_PRESHADED)
        fail(DVR_ERR_UNSUPPORTED, "volume_filter_mode %d is not supported (Nearest=0, Trilinear=1, Preshaded=3)",
             in->volume_filter_mode);
    // renderer.cpp:72-80
    c.preshaded = in->volume_filter_mode == DVR_FILTER_PRESHADED;
    if (c.preshaded != (in->tf_mode == DVR_TF_PRESHADED))
        fail(DVR_ERR_INVALID_ARGUMENT,
             "if pre-shading is selected, both the filter mode and TF mode must be set to preshaded");
    if (c.preshaded && vol.sizes[0] != 4)
        fail(DVR_ERR_INVALID_ARGUMENT,
             "In pre-shaded mode, the volume colors are encoded in the batch dimension of the volume. It must be 4, "
             "but is %lld", (long long)vol.sizes[0]);
    c.B = c.preshaded ? 1 : (int)vol.sizes[0];
    c.X = (int)vol.sizes[1];
    c.Y = (int)vol.sizes[2];
    c.Z = (int)vol.sizes[3];
    check_offsets(vol, "volume");
    check_slice_offsets(vol, "volume");
    P.vol = (const float*)vol.data;
    P.X = c.X; P.Y = c.Y; P.Z = c.Z;
    P.vsb = bstride(vol, 0); P.vsx = bstride(vol, 1); P.vsy = bstride(vol, 2); P.vsz = bstride(vol, 3);

    auto box = [&](const dvr_tensor& t, const float* value, const char* name, const float*& ptr, int& sb, int& sc,
                   float* u) {
        if (given(t)) {
            check_dim(t, name, 2);
            check_device(t, name, c.device);
            check_size(t, name, 1, 3);
            c.B = check_batch(t, name, c.B);
            check_dtype(t, name, DVR_FLOAT32);
            check_offsets(t, name);
            ptr = (const float*)t.data;
            sb = bstride(t, 0);
            sc = (int)t.strides[1];
        } else {
            ptr = nullptr;
            u[0] = value[0]; u[1] = value[1]; u[2] = value[2];
        }
    };
    box(in->box_min, in->box_min_value, "boxMin", P.bmin, P.bminSb, P.bminSc, P.bminU);
    box(in->box_size, in->box_size_value, "boxSize", P.bsize, P.bsizeSb, P.bsizeSc, P.bsizeU);

    c.camera = in->camera_mode;
    P.camMode = c.camera;
    switch (c.camera) {
    case DVR_CAMERA_RAY_START_DIR: {
        const dvr_tensor& s = in->camera_ray_start;
        const dvr_tensor& d = in->camera_ray_dir;
        check_dim(s, "cameraRayStart", 4); check_device(s, "cameraRayStart", c.device);
        check_size(s, "cameraRayStart", 1, c.H); check_size(s, "cameraRayStart", 2, c.W);
        check_size(s, "cameraRayStart", 3, 3);
        c.B = check_batch(s, "cameraRayStart", c.B); check_dtype(s, "cameraRayStart", DVR_FLOAT32);
        check_dim(d, "cameraRayDir", 4); check_device(d, "cameraRayDir", c.device);
        check_size(d, "cameraRayDir", 1, c.H); check_size(d, "cameraRayDir", 2, c.W);
        check_size(d, "cameraRayDir", 3, 3);
        c.B = check_batch(d, "cameraRayDir", c.B); check_dtype(d, "cameraRayDir", DVR_FLOAT32);
        check_offsets(s, "cameraRayStart"); check_offsets(d, "cameraRayDir");
        P.rs = (const float*)s.data;
        P.rsSb = bstride(s, 0); P.rsSy = bstride(s, 1); P.rsSx = bstride(s, 2); P.rsSc = (int)s.strides[3];
        P.rd = (const float*)d.data;
        P.rdSb = bstride(d, 0); P.rdSy = bstride(d, 1); P.rdSx = bstride(d, 2); P.rdSc = (int)d.strides[3];
    } break;
    case DVR_CAMERA_INVERSE_VIEW_MATRIX: {
        const dvr_tensor& m = in->camera_matrix;
        if (given(m)) {
            check_dim(m, "cameraMatrix", 3); check_device(m, "cameraMatrix", c.device);
            check_size(m, "cameraMatrix", 1, 4); check_size(m, "cameraMatrix", 2, 4);
            c.B = check_batch(m, "cameraMatrix", c.B); check_dtype(m, "cameraMatrix", DVR_FLOAT32);
            check_offsets(m, "cameraMatrix");
            P.cm = (const float*)m.data;
            P.cmSb = bstride(m, 0); P.cmSr = (int)m.strides[1]; P.cmSc = (int)m.strides[2];
        } else {
            P.cm = nullptr;
            for (int k = 0; k < 16; ++k) P.cmU[k] = in->camera_matrix_value[k];
        }
    } break;
    case DVR_CAMERA_REFERENCE_FRAME: {
        const dvr_tensor& m = in->camera_matrix;
        check_dim(m, "cameraMatrix", 3); check_device(m, "cameraMatrix", c.device);
        check_size(m, "cameraMatrix", 1, 3); check_size(m, "cameraMatrix", 2, 3);
        c.B = check_batch(m, "cameraMatrix", c.B); check_dtype(m, "cameraMatrix", DVR_FLOAT32);
        check_offsets(m, "cameraMatrix");
        P.cm = (const float*)m.data;
        P.cmSb = bstride(m, 0); P.cmSr = (int)m.strides[1]; P.cmSc = (int)m.strides[2];
        P.fovY = in->camera_fov_y_radians;
        if (c.H > 0) {
            const double fovX = double(P.fovY) * double(c.W) / double(c.H);
            P.tanX = std::tan(fovX / 2.0);
            P.tanY = std::tan(double(P.fovY) / 2.0);
        }
    } break;
    default:
        fail(DVR_ERR_INVALID_ARGUMENT, "unknown camera enum value");
    }

    const dvr_tensor& st = in->step_size;
    if (given(st)) {
        check_dim(st, "stepSize", 3); check_device(st, "stepSize", c.device);
        c.B = check_batch(st, "stepSize", c.B);
        check_batch(st, "stepSize-Height", c.H, 1);
        check_batch(st, "stepSize-Width", c.W, 2);
        check_dtype(st, "stepSize", DVR_FLOAT32);
        check_offsets(st, "stepSize");
        P.st = (const float*)st.data;
        P.stSb = bstride(st, 0); P.stSy = bstride(st, 1); P.stSx = bstride(st, 2);
    } else {
        P.st = nullptr;
        P.stU = in->step_size_value;
        if (!(P.stU > 0.0f) || !std::isfinite(P.stU)) fail(DVR_ERR_INVALID_ARGUMENT, "step size must be positive and finite");
    }

    const dvr_tensor& tf = in->tf;
    check_dim(tf, "tf", 3); check_device(tf, "tf", c.device);
    c.B = check_batch(tf, "tf", c.B);
    check_dtype(tf, "tf", DVR_FLOAT32);
    c.tfm = in->tf_mode;
    switch (c.tfm) {
    case DVR_TF_IDENTITY: check_size(tf, "tf", 1, 1); check_size(tf, "tf", 2, 2); break;
    case DVR_TF_TEXTURE: check_size(tf, "tf", 2, 4); break;
    case DVR_TF_LINEAR:
        check_size(tf, "tf", 2, 5);
        if (tf.sizes[1] <= 1) fail(DVR_ERR_INVALID_ARGUMENT, "tensor 'tf' must have at least two control points in TFLinear-mode");
        break;
    case DVR_TF_GAUSSIAN: check_size(tf, "tf", 2, 6); break;
    case DVR_TF_PRESHADED: break;  // any size, the TF is not read
    default: fail(DVR_ERR_INVALID_ARGUMENT, "unknown tf enum value (tf mode %d)", c.tfm);
    }
    if (tf.sizes[1] < 1) fail(DVR_ERR_INVALID_ARGUMENT, "tensor 'tf' must have at least one row");
    check_offsets(tf, "tf");
    P.tf = (const float*)tf.data;
    P.R = (int)tf.sizes[1]; P.C = (int)tf.sizes[2];
    P.tfSb = bstride(tf, 0); P.tfSr = (int)tf.strides[1]; P.tfSc = (int)tf.strides[2];

    c.filter = in->volume_filter_mode;
    c.blend = in->blend_mode;
    if (c.blend != DVR_BLEND_BEER_LAMBERT && c.blend != DVR_BLEND_ALPHA)
        fail(DVR_ERR_INVALID_ARGUMENT, "unknown blend enum value");
    P.earlyOut = in->blending_early_out;
    if (c.X < 2 || c.Y < 2 || c.Z < 2)
        fail(DVR_ERR_INVALID_ARGUMENT, "volume must have at least 2 voxels per dimension, got %dx%dx%d", c.X, c.Y, c.Z);
    P.W = c.W; P.H = c.H; P.B = c.B;
    if ((int64_t)c.W * c.H * c.B >= (int64_t(1) << 31)) fail(DVR_ERR_INVALID_ARGUMENT, "B*H*W too large");
    return c;
}

void check_outputs(const dvr_outputs* out, Checked& c)
{
    if (!out) fail(DVR_ERR_INVALID_ARGUMENT, "outputs is NULL");
    const dvr_tensor& col = out->color;
    check_dim(col, "color", 4); check_device(col, "color", c.device);
    check_size(col, "color", 0, c.B); check_size(col, "color", 1, c.H); check_size(col, "color", 2, c.W);
    check_size(col, "color", 3, 4); check_dtype(col, "color", DVR_FLOAT32);
    check_offsets(col, "color");
    const dvr_tensor& ti = out->termination_index;
    check_dim(ti, "terminationIndex", 3); check_device(ti, "terminationIndex", c.device);
    check_size(ti, "terminationIndex", 0, c.B); check_size(ti, "terminationIndex", 1, c.H);
    check_size(ti, "terminationIndex", 2, c.W); check_dtype(ti, "terminationIndex", DVR_INT32);
    check_offsets(ti, "terminationIndex");
    c.P.color = (float*)col.data;
    c.P.colSb = (int)col.strides[0]; c.P.colSy = (int)col.strides[1]; c.P.colSx = (int)col.strides[2];
    c.P.colSc = (int)col.strides[3];
    c.P.term = (int*)ti.data;
    c.P.termSb = (int)ti.strides[0]; c.P.termSy = (int)ti.strides[1]; c.P.termSx = (int)ti.strides[2];
}

dvr::AdjointParams check_adjoint(const Checked& c, const dvr_outputs* fwd, const dvr_tensor* adj_color,
                                 const dvr_adjoint_outputs* ao)
{
    dvr::AdjointParams A;
    std::memset(&A, 0, sizeof(A));
    if (!fwd) fail(DVR_ERR_INVALID_ARGUMENT, "outputs_from_forward is NULL");
    const dvr_tensor& col = fwd->color;
    check_dim(col, "output_color", 4); check_device(col, "output_color", c.device);
    check_size(col, "output_color", 0, c.B); check_size(col, "output_color", 1, c.H);
    check_size(col, "output_color", 2, c.W); check_size(col, "output_color", 3, 4);
    check_dtype(col, "output_color", DVR_FLOAT32); check_offsets(col, "output_color");
    const dvr_tensor& ti = fwd->termination_index;
    check_dim(ti, "terminationIndex", 3); check_device(ti, "terminationIndex", c.device);
    check_size(ti, "terminationIndex", 0, c.B); check_size(ti, "terminationIndex", 1, c.H);
    check_size(ti, "terminationIndex", 2, c.W); check_dtype(ti, "terminationIndex", DVR_INT32);
    check_offsets(ti, "terminationIndex");
    A.fcol = (const float*)col.data;
    A.fcSb = (int)col.strides[0]; A.fcSy = (int)col.strides[1]; A.fcSx = (int)col.strides[2]; A.fcSc = (int)col.strides[3];
    A.fterm = (const int*)ti.data;
    A.ftSb = (int)ti.strides[0]; A.ftSy = (int)ti.strides[1]; A.ftSx = (int)ti.strides[2];

    if (!adj_color) fail(DVR_ERR_INVALID_ARGUMENT, "adj_color is NULL");
    const dvr_tensor& ac = *adj_color;
    check_dim(ac, "adj_color", 4); check_device(ac, "adj_color", c.device);
    check_size(ac, "adj_color", 0, c.B); check_size(ac, "adj_color", 1, c.H);
    check_size(ac, "adj_color", 2, c.W); check_size(ac, "adj_color", 3, 4);
    check_dtype(ac, "adj_color", DVR_FLOAT32); check_offsets(ac, "adj_color");
    A.acol = (const float*)ac.data;
    A.acSb = (int)ac.strides[0]; A.acSy = (int)ac.strides[1]; A.acSx = (int)ac.strides[2]; A.acSc = (int)ac.strides[3];

    if (!ao) fail(DVR_ERR_INVALID_ARGUMENT, "adj_outputs is NULL");
    if (ao->has_volume_derivatives) {
        const dvr_tensor& t = ao->adj_volume;
        check_dim(t, "adj_volume", 4); check_device(t, "adj_volume", c.device);
        if (c.preshaded) check_size(t, "adj_volume", 0, 4);  // colour channels (renderer.cpp:354-360)
        else check_batch(t, "adj_volume", c.B);
        check_size(t, "adj_volume", 1, c.X); check_size(t, "adj_volume", 2, c.Y); check_size(t, "adj_volume", 3, c.Z);
        check_dtype(t, "adj_volume", DVR_FLOAT32); check_offsets(t, "adj_volume");
        check_slice_offsets(t, "adj_volume");
        A.avol = (float*)t.data;
        A.avSb = bstride(t, 0); A.avSx = bstride(t, 1); A.avSy = bstride(t, 2); A.avSz = bstride(t, 3);
    }
    if (ao->has_stepsize_derivatives) {
        const dvr_tensor& t = ao->adj_stepsize;
        check_dim(t, "adj_stepSize", 3); check_device(t, "adj_stepSize", c.device);
        check_batch(t, "adj_stepSize", c.B);
        check_batch(t, "adj_stepSize", c.H, 1);
        check_batch(t, "adj_stepSize", c.W, 2);
        check_dtype(t, "adj_stepSize", DVR_FLOAT32); check_offsets(t, "adj_stepSize");
        A.ast = (float*)t.data;
        A.astSb = bstride(t, 0); A.astSy = bstride(t, 1); A.astSx = bstride(t, 2);
        A.stepBroadcast = (c.B != t.sizes[0] || c.H != t.sizes[1] || c.W != t.sizes[2]) ? 1 : 0;
    }
    if (ao->has_camera_derivatives) {
        const dvr_tensor& s = ao->adj_camera_ray_start;
        const dvr_tensor& d = ao->adj_camera_ray_dir;
        check_dim(s, "adj_cameraRayStart", 4); check_device(s, "adj_cameraRayStart", c.device);
        check_batch(s, "adj_cameraRayStart", c.B);
        check_size(s, "adj_cameraRayStart", 1, c.H); check_size(s, "adj_cameraRayStart", 2, c.W);
        check_size(s, "adj_cameraRayStart", 3, 3); check_dtype(s, "adj_cameraRayStart", DVR_FLOAT32);
        check_dim(d, "adj_cameraRayDir", 4); check_device(d, "adj_cameraRayDir", c.device);
        check_batch(d, "adj_cameraRayDir", c.B);
        check_size(d, "adj_cameraRayDir", 1, c.H); check_size(d, "adj_cameraRayDir", 2, c.W);
        check_size(d, "adj_cameraRayDir", 3, 3); check_dtype(d, "adj_cameraRayDir", DVR_FLOAT32);
        check_offsets(s, "adj_cameraRayStart"); check_offsets(d, "adj_cameraRayDir");
        A.ars = (float*)s.data;
        A.arsSb = bstride(s, 0); A.arsSy = bstride(s, 1); A.arsSx = bstride(s, 2); A.arsSc = (int)s.strides[3];
        A.ard = (float*)d.data;
        A.ardSb = bstride(d, 0); A.ardSy = bstride(d, 1); A.ardSx = bstride(d, 2); A.ardSc = (int)d.strides[3];
        A.camBroadcast = (c.B != s.sizes[0] || c.B != d.sizes[0]) ? 1 : 0;
    }
    if (ao->has_tf_derivatives) {
        const dvr_tensor& t = ao->adj_tf;
        check_dim(t, "adj_tf", 3); check_device(t, "adj_tf", c.device);
        check_batch(t, "adj_tf", c.B);
        check_size(t, "adj_tf", 1, c.P.R); check_size(t, "adj_tf", 2, c.P.C);
        check_dtype(t, "adj_tf", DVR_FLOAT32); check_offsets(t, "adj_tf");
        A.atf = (float*)t.data;
        A.atfSb = bstride(t, 0); A.atfSr = (int)t.strides[1]; A.atfSc = (int)t.strides[2];
        A.tfDelayed = ao->tf_delayed_accumulation ? 1 : 0;
    }
    if (!ao->has_tf_derivatives && !ao->has_camera_derivatives && !ao->has_stepsize_derivatives &&
        !ao->has_volume_derivatives)
        fail(DVR_ERR_NO_GRADIENT, "At least one adjoint variable must be activated!");
    return A;
}

// ---------------------------------------------------------------------------
// workspace: grow-only device buffers per device (slot 0: checkpoint and other
// scratch, slot 1: the density store of the fused / replaying adjoint)
// ---------------------------------------------------------------------------
struct Workspace {
    void* ptr = nullptr;
    size_t bytes = 0;
};
constexpr int kWsSlots = 2;
std::mutex g_ws_mutex;
std::vector<Workspace> g_ws;

void* workspace(int device, size_t bytes, hipStream_t stream, int slot = 0)
{
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    if ((int)g_ws.size() <= device * kWsSlots + slot) g_ws.resize((device + 1) * kWsSlots);
    Workspace& w = g_ws[device * kWsSlots + slot];
    if (w.bytes >= bytes) return w.ptr;
    if (w.ptr) {
        // earlier work on this stream may still read the old buffer
        check_hip(hipStreamSynchronize(stream), "hipStreamSynchronize");
        check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
        check_hip(hipFree(w.ptr), "hipFree");
        w.ptr = nullptr;
        w.bytes = 0;
    }
    const size_t want = std::max(bytes, size_t(1) << 20);
    check_hip(hipMalloc(&w.ptr, want), "hipMalloc(workspace)");
    w.bytes = want;
    return w.ptr;
}

// Orders the uses of a device's library scratch (the workspace slots and the static
// array of the allocation-free entry point) across streams and host threads.  A lease
// holds the device's lock from before the scratch is handed out until the work that
// uses it is enqueued; when the previous use was enqueued on another stream, the
// calling stream first waits for the event recorded after it, and the lease records
// that event again when it ends.  Under HIP graph capture neither happens: a captured
// graph orders its own nodes, and its replays use the scratch in the order of the
// stream they are launched on.
struct ScratchSync {
    std::mutex mu;
    hipEvent_t ev = nullptr;
    hipStream_t last = nullptr;
    bool used = false;
};
std::mutex g_sync_mutex;
std::vector<std::unique_ptr<ScratchSync>> g_sync;

class ScratchLease {
public:
    // with `device` current
    ScratchLease(int device, hipStream_t s) : s_(s)
    {
        {
            std::lock_guard<std::mutex> lk(g_sync_mutex);
            if ((int)g_sync.size() <= device) g_sync.resize(device + 1);
            if (!g_sync[device]) g_sync[device] = std::make_unique<ScratchSync>();
            st_ = g_sync[device].get();
        }
        lk_ = std::unique_lock<std::mutex>(st_->mu);
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        check_hip(hipStreamIsCapturing(s, &cs), "hipStreamIsCapturing");
        capturing_ = cs != hipStreamCaptureStatusNone;
        if (capturing_) return;
        if (!st_->ev) check_hip(hipEventCreateWithFlags(&st_->ev, hipEventDisableTiming), "hipEventCreate");
        if (st_->used && st_->last != s) check_hip(hipStreamWaitEvent(s, st_->ev, 0), "hipStreamWaitEvent");
    }
    ~ScratchLease()
    {
        if (capturing_ || !st_->ev) return;
        if (hipEventRecord(st_->ev, s_) == hipSuccess) {
            st_->last = s_;
            st_->used = true;
        } else {
            (void)hipGetLastError();
            st_->used = false;
        }
    }
    ScratchLease(const ScratchLease&) = delete;
    ScratchLease& operator=(const ScratchLease&) = delete;

private:
    hipStream_t s_;
    ScratchSync* st_ = nullptr;
    std::unique_lock<std::mutex> lk_;
    bool capturing_ = false;
};


static_assert(dvr::kMaxTfLds == DVR_MAX_TILED_TF_ENTRIES, "the LDS TF limit is part of the C-ABI");
bool tiled_ok(const Checked& c) { return !c.preshaded && (int64_t)c.P.R * c.P.C <= dvr::kMaxTfLds; }

void no_preshaded(const Checked& c, const char* what)
{
    if (c.preshaded) fail(DVR_ERR_UNSUPPORTED, "%s is not available for pre-shaded volumes", what);
}

// checkpoints per ray: enough for the longest ray when box and step are
// uniform values, else 1024 samples' worth (longer rays recompute from the
// last checkpoint).
int checkpoint_cap(const Checked& c, int K)
{
    int cap;
    if (!c.P.st && !c.P.bsize) {
        const double d = std::sqrt(double(c.P.bsizeU[0]) * c.P.bsizeU[0] + double(c.P.bsizeU[1]) * c.P.bsizeU[1] +
                                   double(c.P.bsizeU[2]) * c.P.bsizeU[2]);
        const double n = d / double(c.P.stU) + 2.0;
        cap = (int)std::ceil(n / K);
    } else {
        cap = 1024 / K;
    }
    return std::max(1, std::min(cap, 1 << 14));
}

int default_k(const Checked&) { return 16; }

// Density store of the checkpointed adjoint: samples per ray it holds (the
// longest ray when box and step are uniform values, else 1024; longer rays
// re-interpolate), and its size; stores above kMaxDensityStore bytes are not
// used by the library's own scratch.
constexpr double kMaxDensityStore = 2.0 * 1024 * 1024 * 1024;
int density_cap(const Checked& c)
{
    double n = 1024.0;
    if (!c.P.st && !c.P.bsize) {
        const double d = std::sqrt(double(c.P.bsizeU[0]) * c.P.bsizeU[0] + double(c.P.bsizeU[1]) * c.P.bsizeU[1] +
                                   double(c.P.bsizeU[2]) * c.P.bsizeU[2]);
        n = std::ceil(d / double(c.P.stU)) + 2.0;
    }
    return (int)std::max(1.0, std::min(n, double(1 << 24)));
}
int64_t density_elements(const Checked& c)
{
    return (int64_t)dvr::tiled_num_tiles(c.P) * density_cap(c) * dvr::kBlock;
}
// the caller's density buffer, validated (float32, contiguous, on the device, large enough)
float* density_buffer(const dvr_tensor& t, const Checked& c)
{
    int64_t numel = 1;
    for (int d = 0; d < t.ndim; ++d) numel *= t.sizes[d];
    if (t.dtype != DVR_FLOAT32 || t.device != c.device)
        fail(DVR_ERR_INVALID_ARGUMENT, "densities must be a float32 tensor on the render device");
    if (numel < density_elements(c))
        fail(DVR_ERR_INVALID_ARGUMENT, "densities needs %lld elements, has %lld", (long long)density_elements(c),
             (long long)numel);
    if (t.ndim > 0 && t.strides[t.ndim - 1] != 1 && numel > 1)
        fail(DVR_ERR_INVALID_ARGUMENT, "densities must be contiguous");
    return (float*)t.data;
}

// Checkpoint interval I in {8, 16, 32, 64}: register segments of K = min(I, 32)
// samples, a checkpoint every I / K segments.
dvr::CkptParams make_ckpt(const Checked& c, int interval)
{
    if (interval != 8 && interval != 16 && interval != 32 && interval != 64)
        fail(DVR_ERR_INVALID_ARGUMENT, "checkpoint_interval must be 8, 16, 32 or 64");
    dvr::CkptParams CK{nullptr, std::min(interval, 32), checkpoint_cap(c, interval), 1};
    CK.every = interval / CK.K;
    return CK;
}


// Host-side stamps of the checkpoint buffers written by dvr_render_forward_ex: the adjoint that
// reads one must use the same checkpoint interval and image batch, and may read densities only
// from the buffer the same forward wrote (else it would read another layout, or uninitialised
// memory, as valid data).  Keyed by the checkpoint buffer's address; a later forward into the
// same buffer replaces the stamp.
struct CkptStamp {
    int K, B, H, W;
    const void* densities;  // written by that forward, or NULL
};
std::mutex g_stampMu;
std::map<const void*, CkptStamp> g_stamps;
constexpr size_t kMaxStamps = 4096;

void stamp_checkpoints(const void* ck, const CkptStamp& st)
{
    std::lock_guard<std::mutex> lk(g_stampMu);
    if (g_stamps.size() >= kMaxStamps && !g_stamps.count(ck)) g_stamps.erase(g_stamps.begin());
    g_stamps[ck] = st;
}

void check_stamp(const void* ck, int K, const Checked& c, const void* densities)
{
    std::lock_guard<std::mutex> lk(g_stampMu);
    auto it = g_stamps.find(ck);
    if (it == g_stamps.end())
        fail(DVR_ERR_INVALID_ARGUMENT, "checkpoints were not written by dvr_render_forward_ex");
    const CkptStamp& st = it->second;
    if (st.K != K)
        fail(DVR_ERR_INVALID_ARGUMENT, "checkpoints were written with checkpoint_interval %d, read with %d", st.K, K);
    if (st.B != c.B || st.H != c.H || st.W != c.W)
        fail(DVR_ERR_INVALID_ARGUMENT, "checkpoints were written for a %dx%dx%d image batch, read for %dx%dx%d",
             st.B, st.H, st.W, c.B, c.H, c.W);
    if (densities && densities != st.densities)
        fail(DVR_ERR_INVALID_ARGUMENT, "densities were not written by the forward that wrote these checkpoints");
}

void do_forward(const dvr_inputs* inputs, dvr_outputs* outputs, const dvr_forward_options* opt, void* stream)
{
    Checked c = check_inputs(inputs);
    check_outputs(outputs, c);
    int path = opt ? opt->path : DVR_PATH_AUTO;
    const int K = opt ? opt->checkpoint_interval : 0;
    if (path == DVR_PATH_AUTO) path = tiled_ok(c) ? DVR_PATH_TILED : DVR_PATH_PLAIN;
    if (c.preshaded && (path != DVR_PATH_PLAIN || K > 0))
        fail(DVR_ERR_UNSUPPORTED, "pre-shaded volumes render on the plain path only");
    DeviceGuard g(c.device);
    hipStream_t s = (hipStream_t)stream;
    if (path == DVR_PATH_TILED && !tiled_ok(c))
        fail(DVR_ERR_UNSUPPORTED, "tiled path needs R*C <= %d transfer-function entries", dvr::kMaxTfLds);
    if (K > 0) {
        if (path != DVR_PATH_TILED) fail(DVR_ERR_UNSUPPORTED, "checkpoints are written by the tiled forward only");
        make_ckpt(c, K);  // validates K
        const dvr_tensor& ck = opt->checkpoints;
        const int64_t need = (int64_t)dvr::tiled_num_tiles(c.P) * checkpoint_cap(c, K) * dvr::kBlock;
        if (!given(ck) || ck.dtype != DVR_FLOAT32 || ck.device != c.device)
            fail(DVR_ERR_INVALID_ARGUMENT, "checkpoints must be a float32 tensor on the render device");
        int64_t numel = 1;
        for (int d = 0; d < ck.ndim; ++d) numel *= ck.sizes[d];
        if (numel < need) fail(DVR_ERR_INVALID_ARGUMENT, "checkpoints needs %lld elements, has %lld", (long long)need, (long long)numel);
        if (ck.ndim > 0 && ck.strides[ck.ndim - 1] != 1 && numel > 1)
            fail(DVR_ERR_INVALID_ARGUMENT, "checkpoints must be contiguous");
        if (given(opt->densities)) density_buffer(opt->densities, c);
    } else if (opt && given(opt->densities)) {
        fail(DVR_ERR_INVALID_ARGUMENT, "densities are written together with checkpoints (checkpoint_interval > 0)");
    }
    if (c.preshaded) {
        check_hip(dvr::launch_forward_preshaded(c.P, c.blend, s), "forward (pre-shaded) launch");
    } else if (path == DVR_PATH_PLAIN) {
        check_hip(dvr::launch_forward_plain(c.P, c.filter, c.tfm, c.blend, s), "forward (plain) launch");
    } else {
        dvr::CkptParams CK{nullptr, 16, 0, 1};
        if (K > 0) {
            CK = make_ckpt(c, K);
            CK.buf = (float*)opt->checkpoints.data;
            if (given(opt->densities)) {
                CK.ds = density_buffer(opt->densities, c);
                CK.dsCap = density_cap(c);
            }
        }
        check_hip(dvr::launch_forward_tiled(c.P, CK, c.filter, c.tfm, c.blend, s), "forward (tiled) launch");
        if (K > 0) stamp_checkpoints(CK.buf, CkptStamp{K, c.B, c.H, c.W, CK.ds});
    }
}

void do_adjoint(const dvr_inputs* inputs, const dvr_outputs* fwd, const dvr_tensor* adj_color,
                dvr_adjoint_outputs* ao, const dvr_adjoint_options* opt, void* stream)
{
    Checked c = check_inputs(inputs);
    dvr::AdjointParams A = check_adjoint(c, fwd, adj_color, ao);
    DeviceGuard g(c.device);
    hipStream_t s = (hipStream_t)stream;
    int method = opt ? opt->method : DVR_ADJOINT_AUTO;
    int path = opt ? opt->path : DVR_PATH_AUTO;
    if (method == DVR_ADJOINT_AUTO) method = tiled_ok(c) ? DVR_ADJOINT_CHECKPOINT : DVR_ADJOINT_INVERSION;
    if (c.preshaded) {
        if (method != DVR_ADJOINT_INVERSION || path == DVR_PATH_TILED)
            fail(DVR_ERR_UNSUPPORTED, "pre-shaded volumes use the inversion adjoint on the plain path");
        check_hip(dvr::launch_adjoint_preshaded(c.P, A, c.blend, s), "adjoint (pre-shaded) launch");
        return;
    }
    if (method == DVR_ADJOINT_INVERSION) {
        if (path == DVR_PATH_TILED) fail(DVR_ERR_UNSUPPORTED, "the inversion adjoint runs on the plain path only");
        check_hip(dvr::launch_adjoint_plain(c.P, A, c.filter, c.tfm, c.blend, s), "adjoint (plain) launch");
        return;
    }
    if (method != DVR_ADJOINT_CHECKPOINT) fail(DVR_ERR_INVALID_ARGUMENT, "unknown adjoint method %d", method);
    if (path == DVR_PATH_PLAIN) fail(DVR_ERR_UNSUPPORTED, "the checkpointed adjoint runs on the tiled path only");
    if (!tiled_ok(c)) fail(DVR_ERR_UNSUPPORTED, "checkpointed adjoint needs R*C <= %d", dvr::kMaxTfLds);
    int K = opt ? opt->checkpoint_interval : 0;
    if (K == 0) K = default_k(c);
    dvr::CkptParams CK = make_ckpt(c, K);
    const int64_t need = (int64_t)dvr::tiled_num_tiles(c.P) * CK.cap * dvr::kBlock;
    bool fromFwd = false;
    std::optional<ScratchLease> lease;  // the library scratch, until the launch is enqueued
    if (opt && given(opt->checkpoints)) {
        const dvr_tensor& ck = opt->checkpoints;
        int64_t numel = 1;
        for (int d = 0; d < ck.ndim; ++d) numel *= ck.sizes[d];
        if (ck.dtype != DVR_FLOAT32 || ck.device != c.device || numel < need)
            fail(DVR_ERR_INVALID_ARGUMENT, "checkpoints do not match these inputs and checkpoint_interval");
        if (ck.ndim > 0 && ck.strides[ck.ndim - 1] != 1 && numel > 1)
            fail(DVR_ERR_INVALID_ARGUMENT, "checkpoints must be contiguous");
        check_stamp(ck.data, K, c, given(opt->densities) ? opt->densities.data : nullptr);
        CK.buf = (float*)ck.data;
        fromFwd = true;
        if (given(opt->densities)) {  // written by the forward launch that wrote the checkpoints
            CK.ds = density_buffer(opt->densities, c);
            CK.dsCap = density_cap(c);
        }
    } else {
        if (opt && given(opt->densities))
            fail(DVR_ERR_INVALID_ARGUMENT, "densities are read together with the checkpoints of the same forward");
        lease.emplace(c.device, s);
        CK.buf = (float*)workspace(c.device, size_t(need) * sizeof(float), s);
        if (double(density_elements(c)) * sizeof(float) <= kMaxDensityStore) {  // the replaying forward writes them
            CK.ds = (float*)workspace(c.device, size_t(density_elements(c)) * sizeof(float), s, 1);
            CK.dsCap = density_cap(c);
        }
    }
    dvr::LossParams L;
    std::memset(&L, 0, sizeof(L));
    check_hip(dvr::launch_adjoint_tiled(c.P, A, CK, L, false, fromFwd, c.filter, c.tfm, c.blend, s),
              "adjoint (tiled) launch");
}

// fixed: the allocation-free variant's scratch (a getter of the device's static array and its
// size in floats); NULL = the library's grow-only workspace
void do_fused(const dvr_inputs* inputs, dvr_outputs* outputs, const dvr_loss* loss, dvr_adjoint_outputs* ao,
              int K, void* stream, const dvr_host::FixedScratch* fixed = nullptr)
{
    Checked c = check_inputs(inputs);
    no_preshaded(c, "fused rendering");
    check_outputs(outputs, c);
    if (!loss) fail(DVR_ERR_INVALID_ARGUMENT, "loss is NULL");
    if (loss->kind != DVR_LOSS_L1 && loss->kind != DVR_LOSS_L2)
        fail(DVR_ERR_INVALID_ARGUMENT, "loss kind must be DVR_LOSS_L1 or DVR_LOSS_L2");
    const dvr_tensor& tg = loss->target;
    check_dim(tg, "target", 4); check_device(tg, "target", c.device);
    check_size(tg, "target", 0, c.B); check_size(tg, "target", 1, c.H); check_size(tg, "target", 2, c.W);
    check_size(tg, "target", 3, 4); check_dtype(tg, "target", DVR_FLOAT32); check_offsets(tg, "target");
    // the adjoint reads colour/termination/adj_color from nowhere in fused mode;
    // validate its outputs against the forward outputs just checked
    dvr_tensor dummy = outputs->color;
    dvr::AdjointParams A = check_adjoint(c, outputs, &dummy, ao);
    dvr::LossParams L;
    std::memset(&L, 0, sizeof(L));
    L.kind = loss->kind;
    L.target = (const float*)tg.data;
    L.tgSb = (int)tg.strides[0]; L.tgSy = (int)tg.strides[1]; L.tgSx = (int)tg.strides[2]; L.tgSc = (int)tg.strides[3];
    const double n = double(c.B) * c.H * c.W * 4.0;
    L.scale = loss->scale != 0.0f ? loss->scale : (n > 0 ? float(1.0 / n) : 0.0f);
    if (given(loss->loss)) {
        const dvr_tensor& lo = loss->loss;
        check_device(lo, "loss", c.device); check_dtype(lo, "loss", DVR_FLOAT32);
        L.lossOut = (float*)lo.data;
    }
    if (!tiled_ok(c)) fail(DVR_ERR_UNSUPPORTED, "fused rendering needs R*C <= %d", dvr::kMaxTfLds);
    if (K == 0) K = default_k(c);
    dvr::CkptParams CK = make_ckpt(c, K);
    DeviceGuard g(c.device);
    hipStream_t s = (hipStream_t)stream;
    const int64_t need = (int64_t)dvr::tiled_num_tiles(c.P) * CK.cap * dvr::kBlock;
    ScratchLease lease(c.device, s);  // the scratch (either kind), until the launch is enqueued
    if (fixed) {
        // checkpoints in the device's static scratch; no density store (the reverse march
        // re-interpolates: the same densities)
        if (need > fixed->floats)
            fail(DVR_ERR_INVALID_ARGUMENT,
                 "dvr_render_fused_noalloc needs %lld floats of checkpoint scratch for these inputs and "
                 "checkpoint_interval %d, its static scratch holds %lld (use a larger interval or fewer pixels "
                 "per call)", (long long)need, K, (long long)fixed->floats);
        CK.buf = fixed->get(c.device);
    } else {
        CK.buf = (float*)workspace(c.device, size_t(need) * sizeof(float), s);
        if (double(density_elements(c)) * sizeof(float) <= kMaxDensityStore) {
            CK.ds = (float*)workspace(c.device, size_t(density_elements(c)) * sizeof(float), s, 1);
            CK.dsCap = density_cap(c);
        }
    }
    check_hip(dvr::launch_adjoint_tiled(c.P, A, CK, L, true, false, c.filter, c.tfm, c.blend, s), "fused launch");
}

// ---------------------------------------------------------------------------
// forward-mode differentiation (renderer/renderer.cpp:419-658)
// ---------------------------------------------------------------------------

int64_t numel(const dvr_tensor& t)
{
    int64_t n = 1;
    for (int d = 0; d < t.ndim; ++d) n *= t.sizes[d];
    return n;
}

void check_index(int idx, int D, const char* name)
{
    if (idx < DVR_NOT_ENABLED || idx >= D)
        fail(DVR_ERR_INVALID_ARGUMENT, "%s = %d is not a derivative index in [0, D=%d) or -1", name, idx, D);
}

// Scalar indices of the settings: every enabled one lies in [0, D).
void check_fwd_indices(const dvr_forward_differences* fd)
{
    if (!fd) fail(DVR_ERR_INVALID_ARGUMENT, "differences_settings is NULL");
    if (fd->D <= 0) fail(DVR_ERR_INVALID_ARGUMENT, "number of derivatives must be positive");
    check_index(fd->d_stepsize, fd->D, "d_stepsize");
    static const char* sn[3] = {"d_rayStart.x", "d_rayStart.y", "d_rayStart.z"};
    static const char* dn[3] = {"d_rayDir.x", "d_rayDir.y", "d_rayDir.z"};
    for (int c = 0; c < 3; ++c) {
        check_index(fd->d_ray_start[c], fd->D, sn[c]);
        check_index(fd->d_ray_dir[c], fd->D, dn[c]);
    }
}

// d_tf as a (batch, R, C) int32 device tensor shaped like the TF
void check_dtf(const dvr_tensor& t, int device, int64_t Bt, int R, int C)
{
    check_dim(t, "d_tf", 3); check_device(t, "d_tf", device);
    check_size(t, "d_tf", 0, Bt); check_size(t, "d_tf", 1, R); check_size(t, "d_tf", 2, C);
    check_dtype(t, "d_tf", DVR_INT32); check_offsets(t, "d_tf");
}

// Copies d_tf to the host (stream-ordered) and builds the inverse table
// entry[b][d] = r*C + c; validates the indices and that no two entries of one
// batch share an index (a single entry per traced index is what the kernel
// seeds).
std::vector<int> tf_entry_table(const dvr_tensor& t, int D, hipStream_t s)
{
    const int Bt = (int)t.sizes[0], R = (int)t.sizes[1], C = (int)t.sizes[2];
    int64_t span = 1;
    for (int d = 0; d < 3; ++d) span += (t.sizes[d] - 1) * t.strides[d];
    std::vector<int> host(size_t(std::max<int64_t>(span, 1)));
    if (Bt * R * C > 0) {
        check_hip(hipMemcpyAsync(host.data(), t.data, sizeof(int) * size_t(span), hipMemcpyDeviceToHost, s),
                  "hipMemcpyAsync(d_tf)");
        check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
    }
    std::vector<int> entry(size_t(Bt) * D, -1);
    for (int b = 0; b < Bt; ++b)
        for (int r = 0; r < R; ++r)
            for (int c = 0; c < C; ++c) {
                const int d = host[size_t(b * t.strides[0] + r * t.strides[1] + c * t.strides[2])];
                if (d == DVR_NOT_ENABLED) continue;
                if (d < DVR_NOT_ENABLED || d >= D)
                    fail(DVR_ERR_INVALID_ARGUMENT, "d_tf[%d][%d][%d] = %d is not a derivative index in [0, D=%d) or -1",
                         b, r, c, d, D);
                int& e = entry[size_t(b) * D + d];
                if (e >= 0)
                    fail(DVR_ERR_UNSUPPORTED, "d_tf maps two TF entries of batch %d to derivative index %d", b, d);
                e = r * C + c;
            }
    return entry;
}

void do_forward_gradients(const dvr_inputs* inputs, const dvr_forward_differences* fd, dvr_outputs* outputs,
                          dvr_tensor* gout, void* stream)
{
    Checked c = check_inputs(inputs);
    no_preshaded(c, "forward-mode differentiation");
    check_fwd_indices(fd);
    const int D = fd->D;
    // d_tf is read only when TF derivatives are enabled (renderer.cpp:430-447)
    if (fd->has_tf_derivatives) check_dtf(fd->d_tf, c.device, inputs->tf.sizes[0], c.P.R, c.P.C);
    const bool hasStep = fd->d_stepsize >= 0;
    bool hasCam = false;
    for (int k = 0; k < 3; ++k) hasCam = hasCam || fd->d_ray_start[k] >= 0 || fd->d_ray_dir[k] >= 0;
    if (!hasStep && !hasCam && !fd->has_tf_derivatives)
        fail(DVR_ERR_INVALID_ARGUMENT, "at least one derivative must be activated");
    check_outputs(outputs, c);
    if (!gout) fail(DVR_ERR_INVALID_ARGUMENT, "gradients_out is NULL");
    const dvr_tensor& g = *gout;
    check_dim(g, "gradients_out", 5); check_device(g, "gradients_out", c.device);
    check_size(g, "gradients_out", 0, c.B); check_size(g, "gradients_out", 1, c.H);
    check_size(g, "gradients_out", 2, c.W); check_size(g, "gradients_out", 3, D);
    check_size(g, "gradients_out", 4, 4); check_dtype(g, "gradients_out", DVR_FLOAT32);
    check_offsets(g, "gradients_out");
    if (!tiled_ok(c)) fail(DVR_ERR_UNSUPPORTED, "forward derivatives need R*C <= %d", dvr::kMaxTfLds);

    DeviceGuard guard(c.device);
    hipStream_t s = (hipStream_t)stream;
    dvr::FwdGradParams G;
    std::memset(&G, 0, sizeof(G));
    G.D = D;
    G.dStep = fd->d_stepsize;
    for (int k = 0; k < 3; ++k) { G.dStart[k] = fd->d_ray_start[k]; G.dDir[k] = fd->d_ray_dir[k]; }
    G.grad = (float*)g.data;
    G.gSb = (int)g.strides[0]; G.gSy = (int)g.strides[1]; G.gSx = (int)g.strides[2];
    G.gSd = (int)g.strides[3]; G.gSc = (int)g.strides[4];
    std::optional<ScratchLease> lease;  // the TF entry table, until the launch is enqueued
    if (fd->has_tf_derivatives) {
        const std::vector<int> entry = tf_entry_table(fd->d_tf, D, s);
        lease.emplace(c.device, s);
        int* dEntry = (int*)workspace(c.device, entry.size() * sizeof(int), s);
        check_hip(hipMemcpyAsync(dEntry, entry.data(), entry.size() * sizeof(int), hipMemcpyHostToDevice, s),
                  "hipMemcpyAsync(tf entry table)");
        // the host table must outlive the copy
        check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
        G.tfEntry = dEntry;
        G.tfEntrySb = fd->d_tf.sizes[0] == 1 ? 0 : D;
    }
    check_hip(dvr::launch_forward_gradients(c.P, G, c.filter, c.tfm, c.blend, s), "forward gradients launch");
}

void do_fwd_to_adjoint(const dvr_tensor* fvp, const dvr_tensor* acp, const dvr_forward_differences* fd,
                       dvr_adjoint_outputs* ao, void* stream)
{
    if (!fvp) fail(DVR_ERR_INVALID_ARGUMENT, "forward_variables is NULL");
    const dvr_tensor& fv = *fvp;
    check_dim(fv, "forwardVariables", 5);
    const int device = fv.device;
    if (device < 0) fail(DVR_ERR_UNSUPPORTED, "forwardVariables is a host (CPU) tensor: this library runs on the GPU only");
    const int B = (int)fv.sizes[0], H = (int)fv.sizes[1], W = (int)fv.sizes[2], D = (int)fv.sizes[3];
    check_size(fv, "forwardVariables", 4, 4); check_dtype(fv, "forwardVariables", DVR_FLOAT32);
    check_offsets(fv, "forwardVariables");
    if (B > 65535) fail(DVR_ERR_UNSUPPORTED, "at most 65535 images per call (batch is %d)", B);
    if (!acp) fail(DVR_ERR_INVALID_ARGUMENT, "adj_color is NULL");
    const dvr_tensor& ac = *acp;
    check_dim(ac, "gradientOutputColor", 4); check_device(ac, "gradientOutputColor", device);
    check_size(ac, "gradientOutputColor", 0, B); check_size(ac, "gradientOutputColor", 1, H);
    check_size(ac, "gradientOutputColor", 2, W); check_size(ac, "gradientOutputColor", 3, 4);
    check_dtype(ac, "gradientOutputColor", DVR_FLOAT32); check_offsets(ac, "gradientOutputColor");
    check_fwd_indices(fd);
    if (fd->D != D)
        fail(DVR_ERR_INVALID_ARGUMENT, "differences_settings.D = %d, but forwardVariables has %d derivatives", fd->D, D);
    if (!ao) fail(DVR_ERR_INVALID_ARGUMENT, "adj_outputs is NULL");
    if (ao->has_volume_derivatives) fail(DVR_ERR_INVALID_ARGUMENT, "volume derivatives are not supported with forward variables");

    dvr::FwdToAdjParams F;
    std::memset(&F, 0, sizeof(F));
    F.B = B; F.H = H; F.W = W; F.D = D;
    F.fv = (const float*)fv.data;
    F.fSb = (int)fv.strides[0]; F.fSy = (int)fv.strides[1]; F.fSx = (int)fv.strides[2];
    F.fSd = (int)fv.strides[3]; F.fSc = (int)fv.strides[4];
    F.ac = (const float*)ac.data;
    F.aSb = (int)ac.strides[0]; F.aSy = (int)ac.strides[1]; F.aSx = (int)ac.strides[2]; F.aSc = (int)ac.strides[3];
    F.dStep = fd->d_stepsize;
    for (int k = 0; k < 3; ++k) { F.dStart[k] = fd->d_ray_start[k]; F.dDir[k] = fd->d_ray_dir[k]; }
    if (ao->has_stepsize_derivatives && F.dStep >= 0) {
        const dvr_tensor& t = ao->adj_stepsize;
        check_dim(t, "adj_stepSize", 3); check_device(t, "adj_stepSize", device);
        check_batch(t, "adj_stepSize", B); check_batch(t, "adj_stepSize", H, 1); check_batch(t, "adj_stepSize", W, 2);
        if (t.sizes[0] > B || t.sizes[1] > H || t.sizes[2] > W)
            fail(DVR_ERR_INVALID_ARGUMENT, "adj_stepSize is larger than the forward variables");
        check_dtype(t, "adj_stepSize", DVR_FLOAT32); check_offsets(t, "adj_stepSize");
        F.ast = (float*)t.data;
        F.astSb = bstride(t, 0); F.astSy = bstride(t, 1); F.astSx = bstride(t, 2);
        F.astAtomic = (t.sizes[0] != B || t.sizes[1] != H || t.sizes[2] != W) ? 1 : 0;
    }
    bool cam = false;
    for (int k = 0; k < 3; ++k) cam = cam || F.dStart[k] >= 0 || F.dDir[k] >= 0;
    if (ao->has_camera_derivatives && cam) {
        const dvr_tensor& s = ao->adj_camera_ray_start;
        const dvr_tensor& d = ao->adj_camera_ray_dir;
        const char* nm[2] = {"adj_cameraRayStart", "adj_cameraRayDir"};
        const dvr_tensor* ts[2] = {&s, &d};
        for (int i = 0; i < 2; ++i) {
            const dvr_tensor& t = *ts[i];
            check_dim(t, nm[i], 4); check_device(t, nm[i], device);
            check_batch(t, nm[i], B);
            if (t.sizes[0] > B) fail(DVR_ERR_INVALID_ARGUMENT, "%s has a larger batch than the forward variables", nm[i]);
            check_size(t, nm[i], 1, H); check_size(t, nm[i], 2, W); check_size(t, nm[i], 3, 3);
            check_dtype(t, nm[i], DVR_FLOAT32); check_offsets(t, nm[i]);
        }
        F.ars = (float*)s.data;
        F.arsSb = bstride(s, 0); F.arsSy = (int)s.strides[1]; F.arsSx = (int)s.strides[2]; F.arsSc = (int)s.strides[3];
        F.ard = (float*)d.data;
        F.ardSb = bstride(d, 0); F.ardSy = (int)d.strides[1]; F.ardSx = (int)d.strides[2]; F.ardSc = (int)d.strides[3];
        F.camAtomic = (s.sizes[0] != B || d.sizes[0] != B) ? 1 : 0;
    }
    const int* dtf = nullptr;
    int dSb = 0, dSr = 0, dSc = 0, tSb = 0, tSr = 0, tSc = 0, R = 0, C = 0;
    float* atf = nullptr;
    if (ao->has_tf_derivatives && fd->has_tf_derivatives) {
        const dvr_tensor& t = ao->adj_tf;
        check_dim(t, "adj_tf", 3); check_device(t, "adj_tf", device);
        check_batch(t, "adj_tf", B);
        if (t.sizes[0] > B) fail(DVR_ERR_INVALID_ARGUMENT, "adj_tf has a larger batch than the forward variables");
        check_dtype(t, "adj_tf", DVR_FLOAT32); check_offsets(t, "adj_tf");
        R = (int)t.sizes[1]; C = (int)t.sizes[2];
        const dvr_tensor& dt = fd->d_tf;
        check_dim(dt, "d_tf", 3);
        check_batch(dt, "d_tf", B);
        if (dt.sizes[0] > B) fail(DVR_ERR_INVALID_ARGUMENT, "d_tf has a larger batch than the forward variables");
        check_dtf(dt, device, dt.sizes[0], R, C);
        if (D > dvr::kMaxFwdSums)
            fail(DVR_ERR_UNSUPPORTED, "TF contraction supports D <= %d derivatives", dvr::kMaxFwdSums);
        dtf = (const int*)dt.data;
        dSb = bstride(dt, 0); dSr = (int)dt.strides[1]; dSc = (int)dt.strides[2];
        atf = (float*)t.data;
        tSb = bstride(t, 0); tSr = (int)t.strides[1]; tSc = (int)t.strides[2];
    }
    if (!F.ast && !F.ars && !atf) fail(DVR_ERR_NO_GRADIENT, "At least one adjoint variable must be activated!");
    DeviceGuard guard(device);
    hipStream_t s = (hipStream_t)stream;
    std::optional<ScratchLease> lease;  // the per-derivative sums, until the launch is enqueued
    if (atf) {
        lease.emplace(device, s);
        F.sums = (float*)workspace(device, sizeof(float) * size_t(B) * D, s);
    }
    check_hip(dvr::launch_forward_to_adjoint(F, dtf, dSb, dSr, dSc, atf, tSb, tSr, tSc, R, C, s),
              "forward variables to gradients launch");
}

void do_compare(const dvr_tensor* colp, const dvr_tensor* gp, const dvr_tensor* refp, dvr_tensor* dop, dvr_tensor* gop,
                int reduce, void* stream)
{
    if (!colp || !gp || !refp || !dop || !gop) fail(DVR_ERR_INVALID_ARGUMENT, "a tensor argument is NULL");
    const dvr_tensor& col = *colp;
    check_dim(col, "colorInput", 4); check_dtype(col, "colorInput", DVR_FLOAT32);
    const int device = col.device;
    if (device < 0) fail(DVR_ERR_UNSUPPORTED, "colorInput is a host (CPU) tensor: this library runs on the GPU only");
    const int B = (int)col.sizes[0], H = (int)col.sizes[1], W = (int)col.sizes[2];
    check_size(col, "colorInput", 3, 4); check_offsets(col, "colorInput");
    if (B > 65535) fail(DVR_ERR_UNSUPPORTED, "at most 65535 images per call (batch is %d)", B);
    const dvr_tensor& g = *gp;
    check_dim(g, "gradientsInput", 5); check_device(g, "gradientsInput", device);
    check_size(g, "gradientsInput", 0, B); check_size(g, "gradientsInput", 1, H); check_size(g, "gradientsInput", 2, W);
    const int D = (int)g.sizes[3];
    check_size(g, "gradientsInput", 4, 4); check_dtype(g, "gradientsInput", DVR_FLOAT32);
    check_offsets(g, "gradientsInput");
    const dvr_tensor& ref = *refp;
    check_dim(ref, "colorReference", 4); check_device(ref, "colorReference", device);
    check_size(ref, "colorReference", 0, B); check_size(ref, "colorReference", 1, H);
    check_size(ref, "colorReference", 2, W); check_size(ref, "colorReference", 3, 4);
    check_dtype(ref, "colorReference", DVR_FLOAT32); check_offsets(ref, "colorReference");
    const dvr_tensor& dout = *dop;
    const dvr_tensor& go = *gop;
    dvr::CompareParams Q;
    std::memset(&Q, 0, sizeof(Q));
    if (reduce) {
        check_dim(dout, "differenceOut", 1); check_device(dout, "differenceOut", device);
        check_size(dout, "differenceOut", 0, 1); check_dtype(dout, "differenceOut", DVR_FLOAT32);
        check_dim(go, "gradientsOut", 1); check_device(go, "gradientsOut", device);
        check_size(go, "gradientsOut", 0, D); check_dtype(go, "gradientsOut", DVR_FLOAT32);
        check_offsets(go, "gradientsOut");
        Q.qSd = (int)go.strides[0];
    } else {
        check_dim(dout, "differenceOut", 4); check_device(dout, "differenceOut", device);
        check_size(dout, "differenceOut", 0, B); check_size(dout, "differenceOut", 1, H);
        check_size(dout, "differenceOut", 2, W); check_size(dout, "differenceOut", 3, 1);
        check_dtype(dout, "differenceOut", DVR_FLOAT32); check_offsets(dout, "differenceOut");
        check_dim(go, "gradientsOut", 4); check_device(go, "gradientsOut", device);
        check_size(go, "gradientsOut", 0, B); check_size(go, "gradientsOut", 1, H);
        check_size(go, "gradientsOut", 2, W); check_size(go, "gradientsOut", 3, D);
        check_dtype(go, "gradientsOut", DVR_FLOAT32); check_offsets(go, "gradientsOut");
        Q.oSb = (int)dout.strides[0]; Q.oSy = (int)dout.strides[1]; Q.oSx = (int)dout.strides[2];
        Q.qSb = (int)go.strides[0]; Q.qSy = (int)go.strides[1]; Q.qSx = (int)go.strides[2]; Q.qSd = (int)go.strides[3];
    }
    Q.B = B; Q.H = H; Q.W = W; Q.D = D;
    Q.col = (const float*)col.data;
    Q.cSb = (int)col.strides[0]; Q.cSy = (int)col.strides[1]; Q.cSx = (int)col.strides[2]; Q.cSc = (int)col.strides[3];
    Q.gin = (const float*)g.data;
    Q.gSb = (int)g.strides[0]; Q.gSy = (int)g.strides[1]; Q.gSx = (int)g.strides[2];
    Q.gSd = (int)g.strides[3]; Q.gSc = (int)g.strides[4];
    Q.ref = (const float*)ref.data;
    Q.rSb = (int)ref.strides[0]; Q.rSy = (int)ref.strides[1]; Q.rSx = (int)ref.strides[2]; Q.rSc = (int)ref.strides[3];
    Q.dout = (float*)dout.data;
    Q.gout = (float*)go.data;
    Q.reduce = reduce ? 1 : 0;
    const double n = double(B) * H * W;
    Q.norm = n > 0 ? float(1.0 / n) : 0.0f;
    DeviceGuard guard(device);
    check_hip(dvr::launch_compare_to_image(Q, (hipStream_t)stream), "compareToImage launch");
}

}  // namespace

// ---------------------------------------------------------------------------
// exported entry points
// ---------------------------------------------------------------------------
extern "C" {

const char* dvr_last_error(void) { return dvr_host::last_error().c_str(); }
int32_t dvr_use_double_precision(void) { return 0; }
const char* dvr_version(void) { return "dvr-mi355x 0.1 (gfx950)"; }

dvr_status dvr_sync(void)
{
    return guarded([&] { check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize"); });
}

dvr_status dvr_render_forward(const dvr_inputs* inputs, dvr_outputs* outputs, void* stream)
{
    return guarded([&] { do_forward(inputs, outputs, nullptr, stream); });
}

dvr_status dvr_render_forward_ex(const dvr_inputs* inputs, dvr_outputs* outputs, const dvr_forward_options* options,
                                 void* stream)
{
    return guarded([&] { do_forward(inputs, outputs, options, stream); });
}

dvr_status dvr_render_adjoint(const dvr_inputs* inputs, const dvr_outputs* outputs_from_forward,
                              const dvr_tensor* adj_color, dvr_adjoint_outputs* adj_outputs, void* stream)
{
    return guarded([&] { do_adjoint(inputs, outputs_from_forward, adj_color, adj_outputs, nullptr, stream); });
}

dvr_status dvr_render_adjoint_ex(const dvr_inputs* inputs, const dvr_outputs* outputs_from_forward,
                                 const dvr_tensor* adj_color, dvr_adjoint_outputs* adj_outputs,
                                 const dvr_adjoint_options* options, void* stream)
{
    return guarded([&] { do_adjoint(inputs, outputs_from_forward, adj_color, adj_outputs, options, stream); });
}

dvr_status dvr_render_fused(const dvr_inputs* inputs, dvr_outputs* outputs, const dvr_loss* loss,
                            dvr_adjoint_outputs* adj_outputs, int32_t checkpoint_interval, void* stream)
{
    return guarded([&] { do_fused(inputs, outputs, loss, adj_outputs, checkpoint_interval, stream); });
}

dvr_status dvr_render_forward_gradients(const dvr_inputs* inputs, const dvr_forward_differences* settings,
                                        dvr_outputs* outputs, dvr_tensor* gradients_out, void* stream)
{
    return guarded([&] { do_forward_gradients(inputs, settings, outputs, gradients_out, stream); });
}

dvr_status dvr_forward_variables_to_gradients(const dvr_tensor* forward_variables, const dvr_tensor* adj_color,
                                              const dvr_forward_differences* settings,
                                              dvr_adjoint_outputs* adj_outputs, void* stream)
{
    return guarded([&] { do_fwd_to_adjoint(forward_variables, adj_color, settings, adj_outputs, stream); });
}

dvr_status dvr_compare_to_image(const dvr_tensor* color, const dvr_tensor* gradients, const dvr_tensor* reference,
                                dvr_tensor* difference_out, dvr_tensor* gradients_out, int32_t reduce, void* stream)
{
    return guarded([&] { do_compare(color, gradients, reference, difference_out, gradients_out, reduce, stream); });
}

int64_t dvr_checkpoint_buffer_elements(const dvr_inputs* inputs, int32_t checkpoint_interval)
{
    if (checkpoint_interval <= 0) return 0;
    int64_t r = -1;
    guarded([&] {
        Checked c = check_inputs(inputs);
        r = (int64_t)dvr::tiled_num_tiles(c.P) * checkpoint_cap(c, checkpoint_interval) * dvr::kBlock;
    });
    return r;
}

int64_t dvr_density_buffer_elements(const dvr_inputs* inputs)
{
    int64_t r = -1;
    guarded([&] {
        Checked c = check_inputs(inputs);
        r = density_elements(c);
    });
    return r;
}

dvr_status dvr_generate_rays(const dvr_tensor* viewport, float fov_y_radians, int32_t width, int32_t height,
                             dvr_tensor* ray_start, dvr_tensor* ray_dir, void* stream)
{
    return guarded([&] {
        if (!viewport || !ray_start || !ray_dir) fail(DVR_ERR_INVALID_ARGUMENT, "NULL argument");
        const dvr_tensor& v = *viewport;
        check_dim(v, "viewport", 3); check_size(v, "viewport", 1, 3); check_size(v, "viewport", 2, 3);
        check_dtype(v, "viewport", DVR_FLOAT32);
        if (v.device < 0) fail(DVR_ERR_UNSUPPORTED, "viewport is a host tensor: this library runs on the GPU only");
        const int B = (int)v.sizes[0];
        for (dvr_tensor* t : {ray_start, ray_dir}) {
            const char* n = t == ray_start ? "rayStart" : "rayDir";
            check_dim(*t, n, 4); check_device(*t, n, v.device); check_size(*t, n, 0, B);
            check_size(*t, n, 1, height); check_size(*t, n, 2, width); check_size(*t, n, 3, 3);
            check_dtype(*t, n, DVR_FLOAT32); check_offsets(*t, n);
        }
        check_offsets(v, "viewport");
        DeviceGuard g(v.device);
        const int rsS[4] = {(int)ray_start->strides[0], (int)ray_start->strides[1], (int)ray_start->strides[2],
                            (int)ray_start->strides[3]};
        const int rdS[4] = {(int)ray_dir->strides[0], (int)ray_dir->strides[1], (int)ray_dir->strides[2],
                            (int)ray_dir->strides[3]};
        check_hip(dvr::launch_generate_rays((const float*)v.data, (int)v.strides[0], (int)v.strides[1],
                                            (int)v.strides[2], B, height, width, fov_y_radians,
                                            (float*)ray_start->data, rsS, (float*)ray_dir->data, rdS,
                                            (hipStream_t)stream),
                  "generate_rays launch");
    });
}

dvr_status dvr_generate_rays_adjoint(const dvr_tensor* viewport, float fov_y_radians, int32_t width, int32_t height,
                                     const dvr_tensor* adj_ray_start, const dvr_tensor* adj_ray_dir,
                                     dvr_tensor* adj_viewport, void* stream)
{
    return guarded([&] {
        if (!viewport || !adj_ray_start || !adj_ray_dir || !adj_viewport) fail(DVR_ERR_INVALID_ARGUMENT, "NULL argument");
        const dvr_tensor& v = *viewport;
        check_dim(v, "viewport", 3); check_size(v, "viewport", 1, 3); check_size(v, "viewport", 2, 3);
        check_dtype(v, "viewport", DVR_FLOAT32);
        if (v.device < 0) fail(DVR_ERR_UNSUPPORTED, "viewport is a host tensor: this library runs on the GPU only");
        const int B = (int)v.sizes[0];
        for (const dvr_tensor* t : {adj_ray_start, adj_ray_dir}) {
            const char* n = t == adj_ray_start ? "grad_rayStart" : "grad_rayDir";
            check_dim(*t, n, 4); check_device(*t, n, v.device); check_size(*t, n, 0, B);
            check_size(*t, n, 1, height); check_size(*t, n, 2, width); check_size(*t, n, 3, 3);
            check_dtype(*t, n, DVR_FLOAT32); check_offsets(*t, n);
        }
        const dvr_tensor& o = *adj_viewport;
        check_dim(o, "adj_viewport", 3); check_device(o, "adj_viewport", v.device); check_size(o, "adj_viewport", 0, B);
        check_size(o, "adj_viewport", 1, 3); check_size(o, "adj_viewport", 2, 3);
        check_dtype(o, "adj_viewport", DVR_FLOAT32);
        check_offsets(v, "viewport"); check_offsets(o, "adj_viewport");
        DeviceGuard g(v.device);
        const int aS[4] = {(int)adj_ray_start->strides[0], (int)adj_ray_start->strides[1],
                           (int)adj_ray_start->strides[2], (int)adj_ray_start->strides[3]};
        const int dS[4] = {(int)adj_ray_dir->strides[0], (int)adj_ray_dir->strides[1], (int)adj_ray_dir->strides[2],
                           (int)adj_ray_dir->strides[3]};
        check_hip(dvr::launch_generate_rays_adjoint((const float*)v.data, (int)v.strides[0], (int)v.strides[1],
                                                    (int)v.strides[2], B, height, width, fov_y_radians,
                                                    (const float*)adj_ray_start->data, aS,
                                                    (const float*)adj_ray_dir->data, dS, (float*)o.data,
                                                    (int)o.strides[0], (int)o.strides[1], (int)o.strides[2],
                                                    (hipStream_t)stream),
                  "generate_rays_adjoint launch");
    });
}

struct dvr_timer {
    hipEvent_t start, stop;
};

dvr_status dvr_timer_create(dvr_timer** timer)
{
    return guarded([&] {
        if (!timer) fail(DVR_ERR_INVALID_ARGUMENT, "NULL timer");
        dvr_timer* t = new dvr_timer{};
        check_hip(hipEventCreate(&t->start), "hipEventCreate");
        check_hip(hipEventCreate(&t->stop), "hipEventCreate");
        *timer = t;
    });
}
dvr_status dvr_timer_destroy(dvr_timer* t)
{
    return guarded([&] {
        if (!t) return;
        (void)hipEventDestroy(t->start);
        (void)hipEventDestroy(t->stop);
        delete t;
    });
}
dvr_status dvr_timer_start(dvr_timer* t, void* stream)
{
    return guarded([&] { check_hip(hipEventRecord(t->start, (hipStream_t)stream), "hipEventRecord"); });
}
dvr_status dvr_timer_stop(dvr_timer* t, void* stream)
{
    return guarded([&] { check_hip(hipEventRecord(t->stop, (hipStream_t)stream), "hipEventRecord"); });
}
dvr_status dvr_timer_elapsed_ms(dvr_timer* t, float* ms)
{
    return guarded([&] {
        check_hip(hipEventSynchronize(t->stop), "hipEventSynchronize");
        check_hip(hipEventElapsedTime(ms, t->start, t->stop), "hipEventElapsedTime");
    });
}

}  // extern "C"

namespace dvr_host {
dvr_status fused_fixed_scratch(const dvr_inputs* inputs, dvr_outputs* outputs, const dvr_loss* loss,
                               dvr_adjoint_outputs* adj_outputs, int32_t checkpoint_interval, void* stream,
                               const FixedScratch& fixed)
{
    return guarded([&] { do_fused(inputs, outputs, loss, adj_outputs, checkpoint_interval, stream, &fixed); });
}
}  // namespace dvr_host
