// C-ABI of the auxiliary kernels: TF pre-shading and its adjoint, best-fit
// inversion, importance sampling (validation of renderer/tf_utils.cpp:173-334
// and renderer/losses.cpp:144-221, same messages) and the elementwise log
// losses (renderer/losses.cpp:8-142).  Every call validates on the host, then
// launches on the caller's stream; nothing is allocated or copied.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>

#include "../../include/dvr.h"
#include "dvr_host_util.hpp"
#include "dvr_launch.hpp"
#include "dvr_params.hpp"

using dvr_host::check_device;
using dvr_host::check_dim;
using dvr_host::check_dtype;
using dvr_host::check_hip;
using dvr_host::check_size;
using dvr_host::DeviceGuard;
using dvr_host::fail;
using dvr_host::given;
using dvr_host::guarded;

namespace {

const char* tf_enum_name(int m)
{
    switch (m) {
    case DVR_TF_IDENTITY: return "TFIdentity";
    case DVR_TF_TEXTURE: return "TFTexture";
    case DVR_TF_LINEAR: return "TFLinear";
    case DVR_TF_GAUSSIAN: return "TFGaussian";
    case DVR_TF_PRESHADED: return "TFPreshaded";
    case DVR_TF_GAUSSIAN_LOG: return "TFGaussianLog";
    default: return "?";
    }
}

void require_gpu(const dvr_tensor& t, const char* what)
{
    if (t.device < 0) fail(DVR_ERR_UNSUPPORTED, "%s: tensors must be on the GPU (this library runs on the GPU only)", what);
}

// every element offset of t must fit into int32
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

dvr::Strides3 strides3(const dvr_tensor& t)
{
    return dvr::Strides3{(int)t.strides[1], (int)t.strides[2], (int)t.strides[3], (int)t.strides[0]};
}

dvr::FastDiv fast_div(unsigned d)
{
    int l = 0;
    while ((1ull << l) < d) ++l;
    const uint64_t m = ((uint64_t(1) << 32) * ((uint64_t(1) << l) - d)) / d + 1;
    return dvr::FastDiv{(unsigned)m, l};
}

// Loop order: the axis of the smallest stride of `t` (dims 1..3) innermost;
// size-1 axes outermost; ties keep z before y before x.
dvr::VoxelGrid voxel_grid(const dvr_tensor& t)
{
    dvr::VoxelGrid G{};
    G.X = (int)t.sizes[1]; G.Y = (int)t.sizes[2]; G.Z = (int)t.sizes[3];
    int ax[3] = {2, 1, 0};
    auto key = [&](int a) { return t.sizes[1 + a] == 1 ? INT64_MAX : t.strides[1 + a]; };
    std::stable_sort(ax, ax + 3, [&](int a, int b) { return key(a) < key(b); });
    const int n[3] = {G.X, G.Y, G.Z};
    for (int k = 0; k < 3; ++k) {
        G.axis[k] = ax[k];
        G.n[k] = n[ax[k]];
    }
    G.div0 = fast_div((unsigned)std::max(G.n[0], 1));
    G.div1 = fast_div((unsigned)std::max(G.n[1], 1));
    G.total = G.X * G.Y * G.Z;
    return G;
}

// 1 x X x Y x Z or 4 x X x Y x Z float32 volume on the GPU (renderer/tf_utils.cpp:179-186)
void check_volume(const dvr_tensor& v, const char* name, int channels, const char* what)
{
    check_dim(v, name, 4);
    check_dtype(v, name, DVR_FLOAT32);
    require_gpu(v, what);
    check_size(v, name, 0, channels);
    for (int d = 1; d < 4; ++d)
        if (v.sizes[d] < 1) fail(DVR_ERR_INVALID_ARGUMENT, "%s must have at least 1 entry at dimension %d", name, d);
    if (v.sizes[1] * v.sizes[2] * v.sizes[3] >= (int64_t(1) << 30))
        fail(DVR_ERR_UNSUPPORTED, "%s: at most 2^30 voxels", name);
    check_offsets(v, name);
}

void check_same_grid(const dvr_tensor& t, const char* name, const dvr_tensor& ref)
{
    for (int d = 1; d < 4; ++d) check_size(t, name, d, ref.sizes[d]);
}

// TF checks of renderer/tf_utils.cpp:188-211 (and :253-275, losses.cpp:167-185)
dvr::TfTable check_tf(const dvr_tensor& tf, int mode, int device, bool allowGaussianLog, bool namedEnum)
{
    check_dim(tf, "tf", 3);
    check_device(tf, "tf", device);
    check_dtype(tf, "tf", DVR_FLOAT32);
    check_size(tf, "tf", 0, 1);
    switch (mode) {
    case DVR_TF_IDENTITY:
        check_size(tf, "tf", 1, 1);
        check_size(tf, "tf", 2, 2);
        break;
    case DVR_TF_TEXTURE:
        check_size(tf, "tf", 2, 4);
        break;
    case DVR_TF_LINEAR:
        check_size(tf, "tf", 2, 5);
        if (tf.sizes[1] <= 1)
            fail(DVR_ERR_INVALID_ARGUMENT, "tensor 'tf' must have at least two control points in TFLinear-mode");
        break;
    case DVR_TF_GAUSSIAN:
        check_size(tf, "tf", 2, 6);
        break;
    case DVR_TF_GAUSSIAN_LOG:
        if (allowGaussianLog) {
            check_size(tf, "tf", 2, 6);
            break;
        }
        [[fallthrough]];
    default:
        if (namedEnum) fail(DVR_ERR_INVALID_ARGUMENT, "unknown tf enum value: %s", tf_enum_name(mode));
        fail(DVR_ERR_INVALID_ARGUMENT, "unknown tf enum value");
    }
    if (tf.sizes[1] < 1) fail(DVR_ERR_INVALID_ARGUMENT, "tf must have at least one entry at dimension 1");
    if (tf.sizes[1] * tf.sizes[2] > dvr::kMaxAuxTf)
        fail(DVR_ERR_UNSUPPORTED, "tf tables of at most %d entries (R*C) are supported here", dvr::kMaxAuxTf);
    check_offsets(tf, "tf");
    return dvr::TfTable{(const float*)tf.data, (int)tf.sizes[1], (int)tf.sizes[2], (int)tf.strides[1],
                        (int)tf.strides[2]};
}

// both outputs/inputs of an elementwise op: contiguous, same shape, dtype and device
int64_t check_elementwise(const dvr_tensor* const* ts, const char* const* names, int n)
{
    const dvr_tensor& r = *ts[0];
    if (!given(r)) fail(DVR_ERR_INVALID_ARGUMENT, "%s is not defined!", names[0]);
    if (r.dtype != DVR_FLOAT32 && r.dtype != DVR_FLOAT64)
        fail(DVR_ERR_INVALID_ARGUMENT, "%s must be of type torch.float32 or torch.float64", names[0]);
    require_gpu(r, names[0]);
    int64_t count = 1;
    for (int d = 0; d < r.ndim; ++d) count *= r.sizes[d];
    for (int k = 0; k < n; ++k) {
        const dvr_tensor& t = *ts[k];
        if (!given(t)) fail(DVR_ERR_INVALID_ARGUMENT, "%s is not defined!", names[k]);
        bool same = t.ndim == r.ndim;
        for (int d = 0; same && d < r.ndim; ++d) same = t.sizes[d] == r.sizes[d];
        if (!same) fail(DVR_ERR_INVALID_ARGUMENT, "both args must have the same shape (%s)", names[k]);
        if (t.device != r.device) fail(DVR_ERR_INVALID_ARGUMENT, "both args must be on the same device (%s)", names[k]);
        if (t.dtype != r.dtype) fail(DVR_ERR_INVALID_ARGUMENT, "both args must have the same dtype (%s)", names[k]);
        int64_t expect = 1;
        for (int d = t.ndim - 1; d >= 0; --d) {
            if (t.sizes[d] != 1 && t.strides[d] != expect)
                fail(DVR_ERR_INVALID_ARGUMENT, "%s must be contiguous", names[k]);
            expect *= t.sizes[d];
        }
    }
    return count;
}

void run_elementwise(int op, const dvr_tensor* const* ts, const char* const* names, int n, int nIn, void* stream)
{
    const int64_t count = check_elementwise(ts, names, n);
    dvr::ElementwiseParams P{};
    const void* in[3] = {nullptr, nullptr, nullptr};
    for (int k = 0; k < nIn; ++k) in[k] = ts[k]->data;
    // inputs: (t) / (t, grad) / (x, y) / (x, y, grad); outputs follow
    if (op == 1) {
        P.a = in[0]; P.g = in[1];
    } else {
        P.a = in[0]; P.b = in[1]; P.g = in[2];
    }
    P.o1 = ts[nIn]->data;
    P.o2 = n > nIn + 1 ? ts[nIn + 1]->data : nullptr;
    P.n = count;
    DeviceGuard g(ts[0]->device);
    check_hip(dvr::launch_elementwise(P, op, ts[0]->dtype == DVR_FLOAT64, (hipStream_t)stream), "elementwise launch");
}

}  // namespace

extern "C" {

dvr_status dvr_preshade_volume(const dvr_tensor* density, const dvr_tensor* tf, int32_t tf_mode, dvr_tensor* color,
                               void* stream)
{
    return guarded([&] {
        if (!density || !tf || !color) fail(DVR_ERR_INVALID_ARGUMENT, "NULL argument");
        check_volume(*density, "densityVolume", 1, "preshade_volume");
        dvr::PreshadeParams P{};
        P.tf = check_tf(*tf, tf_mode, density->device, true, true);
        check_dim(*color, "color", 4);
        check_dtype(*color, "color", DVR_FLOAT32);
        check_device(*color, "color", density->device);
        check_size(*color, "color", 0, 4);
        check_same_grid(*color, "color", *density);
        check_offsets(*color, "color");
        P.G = voxel_grid(*density);
        P.dens = (const float*)density->data; P.ds = strides3(*density);
        P.col = (float*)color->data; P.cs = strides3(*color);
        DeviceGuard g(density->device);
        check_hip(dvr::launch_preshade(P, tf_mode, false, (hipStream_t)stream), "preshade launch");
    });
}

dvr_status dvr_preshade_volume_adjoint(const dvr_tensor* density, const dvr_tensor* tf, int32_t tf_mode,
                                       const dvr_tensor* adj_color, dvr_tensor* adj_density, void* stream)
{
    return guarded([&] {
        if (!density || !tf || !adj_color || !adj_density) fail(DVR_ERR_INVALID_ARGUMENT, "NULL argument");
        check_volume(*density, "densityVolume", 1, "preshade_volume");
        dvr::PreshadeParams P{};
        P.tf = check_tf(*tf, tf_mode, density->device, true, true);
        check_dim(*adj_color, "grad_output", 4);
        check_dtype(*adj_color, "grad_output", DVR_FLOAT32);
        check_device(*adj_color, "grad_output", density->device);
        check_size(*adj_color, "grad_output", 0, 4);
        check_same_grid(*adj_color, "grad_output", *density);
        check_offsets(*adj_color, "grad_output");
        check_dim(*adj_density, "adj_volume", 4);
        check_dtype(*adj_density, "adj_volume", DVR_FLOAT32);
        check_device(*adj_density, "adj_volume", density->device);
        check_size(*adj_density, "adj_volume", 0, 1);
        check_same_grid(*adj_density, "adj_volume", *density);
        check_offsets(*adj_density, "adj_volume");
        P.G = voxel_grid(*density);
        P.dens = (const float*)density->data; P.ds = strides3(*density);
        P.acol = (const float*)adj_color->data; P.as = strides3(*adj_color);
        P.adens = (float*)adj_density->data; P.ads = strides3(*adj_density);
        DeviceGuard g(density->device);
        check_hip(dvr::launch_preshade(P, tf_mode, true, (hipStream_t)stream), "preshade adjoint launch");
    });
}

dvr_status dvr_find_best_fit(const dvr_tensor* color_volume, const dvr_tensor* tf, int32_t tf_mode,
                             int32_t num_samples, float opacity_weighting, const dvr_tensor* previous_densities,
                             float neighbor_weighting, uint64_t seed, dvr_tensor* density, void* stream)
{
    return guarded([&] {
        if (!color_volume || !tf || !density) fail(DVR_ERR_INVALID_ARGUMENT, "NULL argument");
        if (!(opacity_weighting >= 0))
            fail(DVR_ERR_INVALID_ARGUMENT, "opacity weighting must be >= 0, but is %g", (double)opacity_weighting);
        if (num_samples < 1) fail(DVR_ERR_INVALID_ARGUMENT, "numSamples must be >= 1, but is %d", num_samples);
        check_volume(*color_volume, "colorVolume", 4, "find_best_fit");
        dvr::BestFitParams P{};
        P.tf = check_tf(*tf, tf_mode, color_volume->device, false, true);
        const bool prev = previous_densities && given(*previous_densities);
        if (prev) {
            const dvr_tensor& pd = *previous_densities;
            check_dim(pd, "pd", 4);
            check_device(pd, "pd", color_volume->device);
            check_dtype(pd, "pd", DVR_FLOAT32);
            check_size(pd, "pd", 0, 1);
            check_same_grid(pd, "pd", *color_volume);
            check_offsets(pd, "pd");
            P.prev = (const float*)pd.data; P.ps = strides3(pd);
        }
        check_dim(*density, "density", 4);
        check_dtype(*density, "density", DVR_FLOAT32);
        check_device(*density, "density", color_volume->device);
        check_size(*density, "density", 0, 1);
        check_same_grid(*density, "density", *color_volume);
        check_offsets(*density, "density");
        P.G = voxel_grid(*color_volume);
        P.col = (const float*)color_volume->data; P.cs = strides3(*color_volume);
        P.out = (float*)density->data; P.os = strides3(*density);
        P.numSamples = num_samples;
        P.opacityWeighting = opacity_weighting;
        P.neighborWeighting = neighbor_weighting;
        P.seed = seed;
        DeviceGuard g(color_volume->device);
        check_hip(dvr::launch_best_fit(P, tf_mode, (hipStream_t)stream), "best-fit launch");
    });
}

dvr_status dvr_sample_importance(const dvr_tensor* sample_locations, const dvr_tensor* density, const dvr_tensor* tf,
                                 int32_t tf_mode, const float weights[4], int32_t seed, dvr_tensor* maxima,
                                 dvr_tensor* out, void* stream)
{
    return guarded([&] {
        if (!sample_locations || !density || !tf || !weights || !maxima || !out)
            fail(DVR_ERR_INVALID_ARGUMENT, "NULL argument");
        check_dim(*density, "densityVolume", 4);
        check_dtype(*density, "densityVolume", DVR_FLOAT32);
        if (density->device < 0) fail(DVR_ERR_UNSUPPORTED, "sampleImportance is only supported on CUDA");
        check_volume(*density, "densityVolume", 1, "sample_importance");
        dvr::ImportanceParams P{};
        P.tf = check_tf(*tf, tf_mode, density->device, false, false);
        const dvr_tensor& sl = *sample_locations;
        check_dim(sl, "sampleLocations", 2);
        check_device(sl, "sampleLocations", density->device);
        check_dtype(sl, "sampleLocations", DVR_FLOAT32);
        check_size(sl, "sampleLocations", 1, 3);
        check_offsets(sl, "sampleLocations");
        if (sl.sizes[0] >= (int64_t(1) << 30)) fail(DVR_ERR_UNSUPPORTED, "at most 2^30 samples");
        // losses.cpp:188-200: cumulative, normalised weights
        float wU = weights[0], wD = weights[1], wO = weights[2], wG = weights[3];
        if (!(wU >= 0) || !(wD >= 0) || !(wO >= 0) || !(wG >= 0))
            fail(DVR_ERR_INVALID_ARGUMENT, "weights must be non-negative");
        wD += wU;
        wO += wD;
        wG += wO;
        const float sum = wG;
        if (!(sum > 0)) fail(DVR_ERR_INVALID_ARGUMENT, "at least one weight parameter must be positive");
        wU /= sum;
        wD /= sum;
        wO /= sum;
        check_dim(*out, "out", 1);
        check_dtype(*out, "out", DVR_BOOL);
        check_device(*out, "out", density->device);
        check_size(*out, "out", 0, sl.sizes[0]);
        check_offsets(*out, "out");
        check_dim(*maxima, "maxima", 1);
        check_dtype(*maxima, "maxima", DVR_FLOAT32);
        check_device(*maxima, "maxima", density->device);
        check_size(*maxima, "maxima", 0, 3);
        if (maxima->strides[0] != 1) fail(DVR_ERR_INVALID_ARGUMENT, "maxima must be contiguous");
        P.G = voxel_grid(*density);
        P.dens = (const float*)density->data; P.ds = strides3(*density);
        P.loc = (const float*)sl.data; P.lsN = (int)sl.strides[0]; P.lsC = (int)sl.strides[1];
        P.out = (unsigned char*)out->data; P.os = (int)out->strides[0];
        P.N = (int)sl.sizes[0];
        P.maxima = (float*)maxima->data;
        P.wUniform = wU; P.wDensityGradient = wD; P.wOpacity = wO;
        P.seed = (uint64_t)(uint32_t)seed;
        DeviceGuard g(density->device);
        check_hip(dvr::launch_sample_importance(P, tf_mode, (hipStream_t)stream), "importance sampling launch");
    });
}

dvr_status dvr_mul_log(const dvr_tensor* t, dvr_tensor* out, void* stream)
{
    return guarded([&] {
        if (!t || !out) fail(DVR_ERR_INVALID_ARGUMENT, "NULL argument");
        const dvr_tensor* ts[2] = {t, out};
        const char* names[2] = {"t", "output"};
        run_elementwise(0, ts, names, 2, 1, stream);
    });
}

dvr_status dvr_mul_log_backward(const dvr_tensor* t, const dvr_tensor* grad_out, dvr_tensor* grad_t, void* stream)
{
    return guarded([&] {
        if (!t || !grad_out || !grad_t) fail(DVR_ERR_INVALID_ARGUMENT, "NULL argument");
        const dvr_tensor* ts[3] = {t, grad_out, grad_t};
        const char* names[3] = {"t", "grad_output", "grad_t"};
        run_elementwise(1, ts, names, 3, 2, stream);
    });
}

static dvr_status log_loss(int op, const dvr_tensor* x, const dvr_tensor* y, dvr_tensor* out, void* stream)
{
    return guarded([&] {
        if (!x || !y || !out) fail(DVR_ERR_INVALID_ARGUMENT, "NULL argument");
        const dvr_tensor* ts[3] = {x, y, out};
        const char* names[3] = {"logX", "logY", "output"};
        run_elementwise(op, ts, names, 3, 2, stream);
    });
}

static dvr_status log_loss_backward(int op, const dvr_tensor* x, const dvr_tensor* y, const dvr_tensor* g,
                                    dvr_tensor* gx, dvr_tensor* gy, void* stream)
{
    return guarded([&] {
        if (!x || !y || !g || !gx || !gy) fail(DVR_ERR_INVALID_ARGUMENT, "NULL argument");
        const dvr_tensor* ts[5] = {x, y, g, gx, gy};
        const char* names[5] = {"logX", "logY", "grad_output", "grad_logX", "grad_logY"};
        run_elementwise(op, ts, names, 5, 3, stream);
    });
}

dvr_status dvr_log_mse(const dvr_tensor* log_x, const dvr_tensor* log_y, dvr_tensor* out, void* stream)
{
    return log_loss(2, log_x, log_y, out, stream);
}
dvr_status dvr_log_mse_backward(const dvr_tensor* log_x, const dvr_tensor* log_y, const dvr_tensor* grad_out,
                                dvr_tensor* grad_log_x, dvr_tensor* grad_log_y, void* stream)
{
    return log_loss_backward(3, log_x, log_y, grad_out, grad_log_x, grad_log_y, stream);
}
dvr_status dvr_log_l1(const dvr_tensor* log_x, const dvr_tensor* log_y, dvr_tensor* out, void* stream)
{
    return log_loss(4, log_x, log_y, out, stream);
}
dvr_status dvr_log_l1_backward(const dvr_tensor* log_x, const dvr_tensor* log_y, const dvr_tensor* grad_out,
                               dvr_tensor* grad_log_x, dvr_tensor* grad_log_y, void* stream)
{
    return log_loss_backward(5, log_x, log_y, grad_out, grad_log_x, grad_log_y, stream);
}

}  // extern "C"
