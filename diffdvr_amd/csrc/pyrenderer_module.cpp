// pybind11 module `_pyrenderer`: the pyrenderer API surface of DiffDVR
// (bindings/bindings.cpp:132-688) on top of the C-ABI in include/dvr.h.
//
// Tensors are taken duck-typed from PyTorch (data_ptr / shape / stride /
// dtype / device) and handed to libdvr_hip.so as plain strided views; calls
// are issued on the caller's current HIP stream.  This module links only the
// C-ABI library, not libtorch.
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstring>
#include <stdexcept>
#include <string>

#include "dvr.h"

namespace py = pybind11;

namespace {

struct real3 { float x = 0, y = 0, z = 0; };
struct real4 { float x = 0, y = 0, z = 0, w = 0; };
struct int2 { int x = 0, y = 0; };
struct int3 { int x = 0, y = 0, z = 0; };

enum class VolumeFilterMode { Nearest = 0, Trilinear = 1, Preshaded = 3 };
enum class CameraMode { RayStartDir = 0, InverseViewMatrix = 1, ReferenceFrame = 2 };
enum class TFMode { Identity = 0, Texture = 1, Linear = 2, Gaussian = 3, Preshaded = 5, GaussianLog = 6 };
enum class BlendMode { BeerLambert = 0, Alpha = 1 };
enum class Orientation { Xp = 0, Xm = 1, Yp = 2, Ym = 3, Zp = 4, Zm = 5 };

struct CameraPerPixelRays {
    py::object ray_start = py::none(), ray_dir = py::none();
};
struct CameraReferenceFrame {
    py::object viewport = py::none();
    float fov_y_radians = 0;
};
// glm::mat4 stand-in: m[i][j] = buffer[i][j], rows as given (bindings.cpp:330-379)
struct mat4 {
    float m[16] = {0};
};

struct RendererInputs {
    int2 screen_size;
    py::object volume = py::none();
    VolumeFilterMode volume_filter_mode = VolumeFilterMode::Trilinear;
    py::object box_min = py::cast(real3{});
    py::object box_size = py::cast(real3{});
    CameraMode camera_mode = CameraMode::RayStartDir;
    py::object camera = py::none();
    py::object step_size = py::float_(0.0);
    TFMode tf_mode = TFMode::Identity;
    py::object tf = py::none();
    BlendMode blend_mode = BlendMode::BeerLambert;
    float blending_early_out = 1.0f - 1e-5f;
};

struct RendererOutputs {
    py::object color = py::none(), terminationIndex = py::none();
};

struct AdjointOutputs {
    bool has_volume_derivatives = false;
    py::object adj_volume = py::none();
    bool has_stepsize_derivatives = false;
    py::object adj_stepsize = py::none();
    bool has_camera_derivatives = false;
    py::object adj_camera_ray_start = py::none(), adj_camera_ray_dir = py::none();
    bool has_tf_derivatives = false;
    bool tf_delayed_accumulation = false;
    py::object adj_tf = py::none();
};

struct ForwardDifferencesSettings {
    int D = 0;
    int d_stepsize = -1;
    int3 d_rayStart{-1, -1, -1}, d_rayDir{-1, -1, -1};
    py::object d_tf = py::none();
    bool has_tf_derivatives = false;
};

// ---------------------------------------------------------------------------
// tensor -> dvr_tensor
// ---------------------------------------------------------------------------
struct TorchInfo {
    py::object torch, cuda_current_stream, cuda_current_device, float32, int32, float64, bool_;
};
TorchInfo& torch_info()
{
    static TorchInfo* t = nullptr;
    if (!t) {
        t = new TorchInfo();
        t->torch = py::module_::import("torch");
        t->cuda_current_stream = t->torch.attr("cuda").attr("current_stream");
        t->cuda_current_device = t->torch.attr("cuda").attr("current_device");
        t->float32 = t->torch.attr("float32");
        t->int32 = t->torch.attr("int32");
        t->float64 = t->torch.attr("float64");
        t->bool_ = t->torch.attr("bool");
    }
    return *t;
}

dvr_tensor to_tensor(const py::handle& h)
{
    dvr_tensor t;
    std::memset(&t, 0, sizeof(t));
    if (h.is_none()) return t;
    if (!py::hasattr(h, "data_ptr") || !py::hasattr(h, "stride"))
        throw std::invalid_argument("expected a torch.Tensor");
    TorchInfo& ti = torch_info();
    py::tuple shape = py::tuple(h.attr("shape"));
    py::tuple stride = py::tuple(h.attr("stride")());
    t.ndim = (int)shape.size();
    if (t.ndim > 6) throw std::invalid_argument("tensors with more than 6 dimensions are not supported");
    for (int d = 0; d < t.ndim; ++d) {
        t.sizes[d] = shape[d].cast<int64_t>();
        t.strides[d] = stride[d].cast<int64_t>();
    }
    t.data = reinterpret_cast<void*>(h.attr("data_ptr")().cast<uintptr_t>());
    py::object dt = h.attr("dtype");
    t.dtype = dt.is(ti.float32) ? DVR_FLOAT32
              : dt.is(ti.int32)   ? DVR_INT32
              : dt.is(ti.float64) ? DVR_FLOAT64
              : dt.is(ti.bool_)   ? DVR_BOOL
                                  : -1;
    py::object dev = h.attr("device");
    const std::string type = py::str(dev.attr("type"));
    if (type == "cuda") {
        py::object idx = dev.attr("index");
        t.device = idx.is_none() ? ti.cuda_current_device().cast<int>() : idx.cast<int>();
    } else {
        t.device = -1;
    }
    return t;
}

void* current_stream(int device)
{
    if (device < 0) return nullptr;
    py::object s = torch_info().cuda_current_stream(device);
    return reinterpret_cast<void*>(s.attr("cuda_stream").cast<uintptr_t>());
}

void check(dvr_status s)
{
    if (s != DVR_OK) throw std::runtime_error(dvr_last_error());
}

// Keeps the Python objects behind a dvr_inputs alive for the call.
struct PackedInputs {
    dvr_inputs in;
};

PackedInputs pack_inputs(const RendererInputs& r)
{
    PackedInputs p;
    dvr_inputs& in = p.in;
    std::memset(&in, 0, sizeof(in));
    in.screen_width = r.screen_size.x;
    in.screen_height = r.screen_size.y;
    in.volume_filter_mode = (int)r.volume_filter_mode;
    in.volume = to_tensor(r.volume);
    auto vec3 = [](const py::object& o, dvr_tensor& t, float* v) {
        if (py::isinstance<real3>(o)) {
            const real3 x = o.cast<real3>();
            v[0] = x.x; v[1] = x.y; v[2] = x.z;
        } else {
            t = to_tensor(o);
        }
    };
    vec3(r.box_min, in.box_min, in.box_min_value);
    vec3(r.box_size, in.box_size, in.box_size_value);
    in.camera_mode = (int)r.camera_mode;
    switch (r.camera_mode) {
    case CameraMode::RayStartDir:
        if (!py::isinstance<CameraPerPixelRays>(r.camera))
            throw std::runtime_error("'CameraRayStartDir' specified in 'cameraMode', but 'camera' does not hold the "
                                     "first variant 'CameraPerPixelRays'");
        {
            const CameraPerPixelRays& c = r.camera.cast<const CameraPerPixelRays&>();
            in.camera_ray_start = to_tensor(c.ray_start);
            in.camera_ray_dir = to_tensor(c.ray_dir);
        }
        break;
    case CameraMode::InverseViewMatrix:
        if (py::isinstance<CameraPerPixelRays>(r.camera) || py::isinstance<CameraReferenceFrame>(r.camera))
            throw std::runtime_error("'CameraInverseViewMatrix' specified in 'cameraMode', but 'camera' holds neither "
                                     "a B*4*4 tensor nor a mat4");
        if (py::isinstance<mat4>(r.camera)) {
            const mat4& m = r.camera.cast<const mat4&>();
            std::memcpy(in.camera_matrix_value, m.m, sizeof(m.m));
        } else {
            in.camera_matrix = to_tensor(r.camera);
        }
        break;
    case CameraMode::ReferenceFrame:
        if (!py::isinstance<CameraReferenceFrame>(r.camera))
            throw std::runtime_error("'CameraReferenceFrame' specified in 'cameraMode', but 'camera' does not hold the "
                                     "fourth variant 'CameraReferenceFrame'");
        {
            const CameraReferenceFrame& c = r.camera.cast<const CameraReferenceFrame&>();
            in.camera_matrix = to_tensor(c.viewport);
            in.camera_fov_y_radians = c.fov_y_radians;
        }
        break;
    }
    if (py::isinstance<py::float_>(r.step_size) || py::isinstance<py::int_>(r.step_size))
        in.step_size_value = r.step_size.cast<float>();
    else
        in.step_size = to_tensor(r.step_size);
    in.tf_mode = (int)r.tf_mode;
    in.tf = to_tensor(r.tf);
    in.blend_mode = (int)r.blend_mode;
    in.blending_early_out = r.blending_early_out;
    return p;
}

dvr_outputs pack_outputs(const RendererOutputs& o)
{
    dvr_outputs out;
    std::memset(&out, 0, sizeof(out));
    out.color = to_tensor(o.color);
    out.termination_index = to_tensor(o.terminationIndex);
    return out;
}

dvr_forward_differences pack_differences(const ForwardDifferencesSettings& f)
{
    dvr_forward_differences d;
    std::memset(&d, 0, sizeof(d));
    d.D = f.D;
    d.d_stepsize = f.d_stepsize;
    d.d_ray_start[0] = f.d_rayStart.x; d.d_ray_start[1] = f.d_rayStart.y; d.d_ray_start[2] = f.d_rayStart.z;
    d.d_ray_dir[0] = f.d_rayDir.x; d.d_ray_dir[1] = f.d_rayDir.y; d.d_ray_dir[2] = f.d_rayDir.z;
    d.has_tf_derivatives = f.has_tf_derivatives ? 1 : 0;
    if (f.has_tf_derivatives) d.d_tf = to_tensor(f.d_tf);
    return d;
}

dvr_adjoint_outputs pack_adjoint(const AdjointOutputs& a)
{
    dvr_adjoint_outputs o;
    std::memset(&o, 0, sizeof(o));
    o.has_volume_derivatives = a.has_volume_derivatives;
    if (a.has_volume_derivatives) o.adj_volume = to_tensor(a.adj_volume);
    o.has_stepsize_derivatives = a.has_stepsize_derivatives;
    if (a.has_stepsize_derivatives) o.adj_stepsize = to_tensor(a.adj_stepsize);
    o.has_camera_derivatives = a.has_camera_derivatives;
    if (a.has_camera_derivatives) {
        o.adj_camera_ray_start = to_tensor(a.adj_camera_ray_start);
        o.adj_camera_ray_dir = to_tensor(a.adj_camera_ray_dir);
    }
    o.has_tf_derivatives = a.has_tf_derivatives;
    if (a.has_tf_derivatives) o.adj_tf = to_tensor(a.adj_tf);
    o.tf_delayed_accumulation = a.tf_delayed_accumulation;
    return o;
}

int parse_path(const std::string& s)
{
    if (s == "auto") return DVR_PATH_AUTO;
    if (s == "plain") return DVR_PATH_PLAIN;
    if (s == "tiled") return DVR_PATH_TILED;
    throw std::invalid_argument("path must be 'auto', 'plain' or 'tiled'");
}
int parse_method(const std::string& s)
{
    if (s == "auto") return DVR_ADJOINT_AUTO;
    if (s == "inversion") return DVR_ADJOINT_INVERSION;
    if (s == "checkpoint") return DVR_ADJOINT_CHECKPOINT;
    throw std::invalid_argument("method must be 'auto', 'inversion' or 'checkpoint'");
}

struct GpuTimer {
    dvr_timer* t = nullptr;
    GpuTimer() { check(dvr_timer_create(&t)); }
    ~GpuTimer() { dvr_timer_destroy(t); }
    void start() { check(dvr_timer_start(t, current_stream(torch_info().cuda_current_device().cast<int>()))); }
    void stop() { check(dvr_timer_stop(t, current_stream(torch_info().cuda_current_device().cast<int>()))); }
    float elapsed_ms()
    {
        float ms = 0;
        check(dvr_timer_elapsed_ms(t, &ms));
        return ms;
    }
};


// ---------------------------------------------------------------------------
// batched plans and the Adam group: thin handles over the C-ABI; tensors are
// validated and kept alive by diffdvr_amd/batch.py, which passes raw pointers.
// ---------------------------------------------------------------------------
struct BatchPlanHandle {
    dvr_batch_plan* plan = nullptr;
    ~BatchPlanHandle()
    {
        if (plan) (void)dvr_batch_plan_destroy(plan);
    }
};

template <class T> T* ptr_of(const py::handle& h) { return reinterpret_cast<T*>(h.cast<uintptr_t>()); }

std::unique_ptr<BatchPlanHandle> make_plan(py::list volumes, py::list tfs, py::list views, int filter, int tfm,
                                           int blend, float early, int loss, uintptr_t loss_out, bool forward_only,
                                           int device, int K, int ww, int grid, bool reinterpolate)
{
    std::vector<dvr_batch_volume> vs;
    for (auto h : volumes) {
        py::tuple t = py::reinterpret_borrow<py::tuple>(h);
        if (t.size() != 7) throw std::invalid_argument("volume entry: (data, X, Y, Z, box_min, box_size, adj)");
        dvr_batch_volume v;
        std::memset(&v, 0, sizeof(v));
        v.data = ptr_of<const float>(t[0]);
        v.size_x = t[1].cast<int>(); v.size_y = t[2].cast<int>(); v.size_z = t[3].cast<int>();
        py::sequence bm = py::reinterpret_borrow<py::sequence>(t[4]), bs = py::reinterpret_borrow<py::sequence>(t[5]);
        for (int k = 0; k < 3; ++k) { v.box_min[k] = bm[k].cast<float>(); v.box_size[k] = bs[k].cast<float>(); }
        v.adj_data = ptr_of<float>(t[6]);
        vs.push_back(v);
    }
    std::vector<dvr_batch_tf> ts;
    for (auto h : tfs) {
        py::tuple t = py::reinterpret_borrow<py::tuple>(h);
        if (t.size() != 4) throw std::invalid_argument("tf entry: (data, rows, cols, adj)");
        dvr_batch_tf f;
        std::memset(&f, 0, sizeof(f));
        f.data = ptr_of<const float>(t[0]);
        f.rows = t[1].cast<int>(); f.cols = t[2].cast<int>();
        f.adj_data = ptr_of<float>(t[3]);
        ts.push_back(f);
    }
    std::vector<dvr_batch_view> ws;
    for (auto h : views) {
        py::tuple t = py::reinterpret_borrow<py::tuple>(h);
        if (t.size() != 14)
            throw std::invalid_argument("view entry: (volume, tf, W, H, camera_mode, camera, fov_y, step, color, "
                                        "termination, target, adj_color, loss_scale, adj_camera)");
        dvr_batch_view w;
        std::memset(&w, 0, sizeof(w));
        w.volume = t[0].cast<int>(); w.tf = t[1].cast<int>();
        w.width = t[2].cast<int>(); w.height = t[3].cast<int>();
        w.camera_mode = t[4].cast<int>();
        w.camera = ptr_of<const float>(t[5]);
        w.fov_y_radians = t[6].cast<float>();
        w.step_size = t[7].cast<float>();
        w.color = ptr_of<float>(t[8]);
        w.termination_index = ptr_of<int32_t>(t[9]);
        w.target = ptr_of<const float>(t[10]);
        w.adj_color = ptr_of<const float>(t[11]);
        w.loss_scale = t[12].cast<float>();
        w.adj_camera = ptr_of<float>(t[13]);
        ws.push_back(w);
    }
    dvr_batch b;
    std::memset(&b, 0, sizeof(b));
    b.num_volumes = (int)vs.size(); b.num_tfs = (int)ts.size(); b.num_views = (int)ws.size();
    b.volumes = vs.data(); b.tfs = ts.data(); b.views = ws.data();
    b.volume_filter_mode = filter; b.tf_mode = tfm; b.blend_mode = blend; b.blending_early_out = early;
    b.loss = loss; b.loss_out = reinterpret_cast<float*>(loss_out); b.forward_only = forward_only ? 1 : 0;
    b.device = device;
    dvr_batch_options o{K, ww, grid, reinterpolate ? 1 : 0};
    std::unique_ptr<BatchPlanHandle> hdl(new BatchPlanHandle());
    check(dvr_batch_plan_create(&b, &o, &hdl->plan));
    return hdl;
}

py::tuple plan_options(const BatchPlanHandle& h)
{
    dvr_batch_options o;
    check(dvr_batch_plan_get_options(h.plan, &o));
    return py::make_tuple(o.checkpoint_interval, o.wave_width, o.grid, o.reinterpolate != 0);
}

struct AdamHandle {
    dvr_adam_group* group = nullptr;
    ~AdamHandle()
    {
        if (group) (void)dvr_adam_group_destroy(group);
    }
};

}  // namespace

PYBIND11_MODULE(_pyrenderer, m)
{
    m.attr("MAX_TILED_TF_ENTRIES") = DVR_MAX_TILED_TF_ENTRIES;
    m.attr("NOALLOC_SCRATCH_FLOATS") = (int64_t)DVR_NOALLOC_SCRATCH_FLOATS;
    m.doc() = "MI355X (gfx950) differentiable direct volume renderer: pyrenderer API over libdvr_hip";
    m.def("use_double_precision", []() { return dvr_use_double_precision(); },
          "Returns 0: this library computes in single precision.");
    m.def("version", []() { return std::string(dvr_version()); });
    m.def("sync", []() { check(dvr_sync()); });
    m.def("cleanup", []() {});
    m.def("reload_kernels", []() {});
    m.def("set_cuda_sync_mode", [](bool) {});
    m.def("set_cuda_debug_mode", [](bool) {});
    m.def("set_cuda_cache_dir", [](const std::string&) {});
    m.def("disable_cuda_cache", []() {});

    py::class_<real3>(m, "real3")
        .def(py::init<>())
        .def(py::init([](float x, float y, float z) { return real3{x, y, z}; }))
        .def_readwrite("x", &real3::x).def_readwrite("y", &real3::y).def_readwrite("z", &real3::z)
        .def("__str__", [](const real3& v) { return "(" + std::to_string(v.x) + ", " + std::to_string(v.y) + ", " + std::to_string(v.z) + ")"; });
    py::class_<real4>(m, "real4")
        .def(py::init<>())
        .def(py::init([](float x, float y, float z, float w) { return real4{x, y, z, w}; }))
        .def_readwrite("x", &real4::x).def_readwrite("y", &real4::y).def_readwrite("z", &real4::z)
        .def_readwrite("w", &real4::w)
        .def("__str__", [](const real4& v) {
            return "(" + std::to_string(v.x) + ", " + std::to_string(v.y) + ", " + std::to_string(v.z) + ", " + std::to_string(v.w) + ")";
        });
    py::class_<int2>(m, "int2")
        .def(py::init<>())
        .def(py::init([](int x, int y) { return int2{x, y}; }))
        .def_readwrite("x", &int2::x).def_readwrite("y", &int2::y)
        .def("__str__", [](const int2& v) { return "(" + std::to_string(v.x) + ", " + std::to_string(v.y) + ")"; });
    py::class_<int3>(m, "int3")
        .def(py::init<>())
        .def(py::init([](int x, int y, int z) { return int3{x, y, z}; }))
        .def_readwrite("x", &int3::x).def_readwrite("y", &int3::y).def_readwrite("z", &int3::z)
        .def("__str__", [](const int3& v) { return "(" + std::to_string(v.x) + ", " + std::to_string(v.y) + ", " + std::to_string(v.z) + ")"; });

    py::enum_<VolumeFilterMode>(m, "VolumeFilterMode")
        .value("Nearest", VolumeFilterMode::Nearest)
        .value("Trilinear", VolumeFilterMode::Trilinear)
        .value("Preshaded", VolumeFilterMode::Preshaded);
    py::enum_<CameraMode>(m, "CameraMode")
        .value("RayStartDir", CameraMode::RayStartDir)
        .value("InverseViewMatrix", CameraMode::InverseViewMatrix)
        .value("ReferenceFrame", CameraMode::ReferenceFrame);
    py::enum_<TFMode>(m, "TFMode")
        .value("Identity", TFMode::Identity)
        .value("Texture", TFMode::Texture)
        .value("Linear", TFMode::Linear)
        .value("Gaussian", TFMode::Gaussian)
        .value("GaussianLog", TFMode::GaussianLog)
        .value("Preshaded", TFMode::Preshaded);
    py::enum_<BlendMode>(m, "BlendMode")
        .value("BeerLambert", BlendMode::BeerLambert)
        .value("Alpha", BlendMode::Alpha);
    py::enum_<Orientation>(m, "Orientation")
        .value("Xp", Orientation::Xp).value("Xm", Orientation::Xm)
        .value("Yp", Orientation::Yp).value("Ym", Orientation::Ym)
        .value("Zp", Orientation::Zp).value("Zm", Orientation::Zm);

    py::class_<CameraPerPixelRays>(m, "CameraPerPixelRays")
        .def(py::init<>())
        .def(py::init([](py::object s, py::object d) { return CameraPerPixelRays{s, d}; }))
        .def_readwrite("ray_start", &CameraPerPixelRays::ray_start, "Tensor of ray starting positions of shape B*H*W*3")
        .def_readwrite("ray_dir", &CameraPerPixelRays::ray_dir, "Tensor of ray direction of shape B*H*W*3");
    py::class_<CameraReferenceFrame>(m, "CameraReferenceFrame")
        .def(py::init<>())
        .def(py::init([](py::object v, float fov) { return CameraReferenceFrame{v, fov}; }))
        .def_readwrite("viewport", &CameraReferenceFrame::viewport,
                       "Tensor of the viewport matrix of shape B*3*3 (eye, right, up)")
        .def_readwrite("fov_y_radians", &CameraReferenceFrame::fov_y_radians, "field-of-view for the y-axis in radians");
    py::class_<mat4>(m, "mat4")
        .def(py::init([](py::object b) {
            mat4 r;
            py::sequence rows = py::reinterpret_borrow<py::sequence>(b.attr("tolist")());
            if (rows.size() != 4) throw std::runtime_error("Incompatible buffer sizes! Expected a shape of 4x4");
            for (int i = 0; i < 4; ++i) {
                py::sequence row = rows[i];
                if (row.size() != 4) throw std::runtime_error("Incompatible buffer sizes! Expected a shape of 4x4");
                for (int j = 0; j < 4; ++j) r.m[4 * i + j] = row[j].cast<float>();
            }
            return r;
        }))
        .def("tolist", [](const mat4& m) {
            py::list l;
            for (int i = 0; i < 4; ++i) {
                py::list row;
                for (int j = 0; j < 4; ++j) row.append(m.m[4 * i + j]);
                l.append(row);
            }
            return l;
        });

    py::class_<RendererInputs>(m, "RendererInputs")
        .def(py::init<>())
        .def("clone", [](const RendererInputs& a) { return RendererInputs(a); }, "performs a shallow clone")
        .def_readwrite("screen_size", &RendererInputs::screen_size, "the screen size (width, height) of type int2")
        .def_readwrite("volume", &RendererInputs::volume, "the volume tensor of shape B*X*Y*Z")
        .def_readwrite("volume_filter_mode", &RendererInputs::volume_filter_mode)
        .def_readwrite("box_min", &RendererInputs::box_min, "real3 or float tensor of shape B*3")
        .def_readwrite("box_size", &RendererInputs::box_size, "real3 or float tensor of shape B*3")
        .def_readwrite("camera_mode", &RendererInputs::camera_mode)
        .def_readwrite("camera", &RendererInputs::camera,
                       "CameraPerPixelRays | B*4*4 tensor or mat4 | CameraReferenceFrame, per camera_mode")
        .def_readwrite("step_size", &RendererInputs::step_size, "float or float tensor of shape B*H*W")
        .def_readwrite("tf_mode", &RendererInputs::tf_mode)
        .def_readwrite("tf", &RendererInputs::tf, "transfer function tensor of shape B*R*C")
        .def_readwrite("blend_mode", &RendererInputs::blend_mode)
        .def_readwrite("blending_early_out", &RendererInputs::blending_early_out);

    py::class_<RendererOutputs>(m, "RendererOutputs")
        .def(py::init<>())
        .def(py::init([](py::object c, py::object t) { return RendererOutputs{c, t}; }))
        .def_readwrite("color", &RendererOutputs::color)
        .def_readwrite("terminationIndex", &RendererOutputs::terminationIndex);

    py::class_<ForwardDifferencesSettings>(m, "ForwardDifferencesSettings")
        .def(py::init<>())
        .def_readwrite("D", &ForwardDifferencesSettings::D)
        .def_readwrite("d_stepsize", &ForwardDifferencesSettings::d_stepsize)
        .def_readwrite("d_rayStart", &ForwardDifferencesSettings::d_rayStart)
        .def_readwrite("d_rayDir", &ForwardDifferencesSettings::d_rayDir)
        .def_readwrite("d_tf", &ForwardDifferencesSettings::d_tf)
        .def_readwrite("has_tf_derivatives", &ForwardDifferencesSettings::has_tf_derivatives);

    py::class_<AdjointOutputs>(m, "AdjointOutputs")
        .def(py::init<>())
        .def_readwrite("has_volume_derivatives", &AdjointOutputs::has_volume_derivatives)
        .def_readwrite("adj_volume", &AdjointOutputs::adj_volume)
        .def_readwrite("has_stepsize_derivatives", &AdjointOutputs::has_stepsize_derivatives)
        .def_readwrite("adj_stepsize", &AdjointOutputs::adj_stepsize)
        .def_readwrite("has_camera_derivatives", &AdjointOutputs::has_camera_derivatives)
        .def_readwrite("adj_camera_ray_start", &AdjointOutputs::adj_camera_ray_start)
        .def_readwrite("adj_camera_ray_dir", &AdjointOutputs::adj_camera_ray_dir)
        .def_readwrite("has_tf_derivatives", &AdjointOutputs::has_tf_derivatives)
        .def_readwrite("tf_delayed_accumulation", &AdjointOutputs::tf_delayed_accumulation)
        .def_readwrite("adj_tf", &AdjointOutputs::adj_tf);

    py::class_<GpuTimer>(m, "GpuTimer")
        .def(py::init<>())
        .def("start", &GpuTimer::start)
        .def("stop", &GpuTimer::stop)
        .def("elapsed_ms", &GpuTimer::elapsed_ms);

    m.def("render_forward",
          [](const RendererInputs& inputs, RendererOutputs& outputs, const std::string& path, int checkpoint_interval,
             py::object checkpoints, py::object densities) {
              PackedInputs p = pack_inputs(inputs);
              dvr_outputs o = pack_outputs(outputs);
              dvr_forward_options opt;
              std::memset(&opt, 0, sizeof(opt));
              opt.path = parse_path(path);
              opt.checkpoint_interval = checkpoint_interval;
              opt.checkpoints = to_tensor(checkpoints);
              opt.densities = to_tensor(densities);
              check(dvr_render_forward_ex(&p.in, &o, &opt, current_stream(p.in.volume.device)));
          },
          py::arg("inputs"), py::arg("outputs"), py::arg("path") = "auto", py::arg("checkpoint_interval") = 0,
          py::arg("checkpoints") = py::none(), py::arg("densities") = py::none());
    m.def("render_adjoint",
          [](const RendererInputs& inputs, const RendererOutputs& fwd, py::object adj_color, AdjointOutputs& adj,
             const std::string& method, const std::string& path, int checkpoint_interval, py::object checkpoints,
             py::object densities) {
              PackedInputs p = pack_inputs(inputs);
              dvr_outputs o = pack_outputs(fwd);
              dvr_tensor ac = to_tensor(adj_color);
              dvr_adjoint_outputs ao = pack_adjoint(adj);
              dvr_adjoint_options opt;
              std::memset(&opt, 0, sizeof(opt));
              opt.method = parse_method(method);
              opt.path = parse_path(path);
              opt.checkpoint_interval = checkpoint_interval;
              opt.checkpoints = to_tensor(checkpoints);
              opt.densities = to_tensor(densities);
              check(dvr_render_adjoint_ex(&p.in, &o, &ac, &ao, &opt, current_stream(p.in.volume.device)));
          },
          py::arg("inputs"), py::arg("outputs_from_forward"), py::arg("adj_color"), py::arg("adj_outputs"),
          py::arg("method") = "auto", py::arg("path") = "auto", py::arg("checkpoint_interval") = 0,
          py::arg("checkpoints") = py::none(), py::arg("densities") = py::none());
    m.def("render_fused",
          [](const RendererInputs& inputs, RendererOutputs& outputs, const std::string& loss, py::object target,
             AdjointOutputs& adj, py::object loss_out, float scale, int checkpoint_interval) {
              PackedInputs p = pack_inputs(inputs);
              dvr_outputs o = pack_outputs(outputs);
              dvr_loss l;
              std::memset(&l, 0, sizeof(l));
              if (loss == "l1") l.kind = DVR_LOSS_L1;
              else if (loss == "l2" || loss == "mse") l.kind = DVR_LOSS_L2;
              else throw std::invalid_argument("loss must be 'l1' or 'l2'");
              l.target = to_tensor(target);
              l.scale = scale;
              l.loss = to_tensor(loss_out);
              dvr_adjoint_outputs ao = pack_adjoint(adj);
              check(dvr_render_fused(&p.in, &o, &l, &ao, checkpoint_interval, current_stream(p.in.volume.device)));
          },
          py::arg("inputs"), py::arg("outputs"), py::arg("loss"), py::arg("target"), py::arg("adj_outputs"),
          py::arg("loss_out") = py::none(), py::arg("scale") = 0.0f, py::arg("checkpoint_interval") = 0);
    m.def("render_fused_noalloc",
          [](const RendererInputs& inputs, RendererOutputs& outputs, const std::string& loss, py::object target,
             AdjointOutputs& adj, py::object loss_out, float scale, int checkpoint_interval) {
              PackedInputs p = pack_inputs(inputs);
              dvr_outputs o = pack_outputs(outputs);
              dvr_loss l;
              std::memset(&l, 0, sizeof(l));
              if (loss == "l1") l.kind = DVR_LOSS_L1;
              else if (loss == "l2" || loss == "mse") l.kind = DVR_LOSS_L2;
              else throw std::invalid_argument("loss must be 'l1' or 'l2'");
              l.target = to_tensor(target);
              l.scale = scale;
              l.loss = to_tensor(loss_out);
              dvr_adjoint_outputs ao = pack_adjoint(adj);
              check(dvr_render_fused_noalloc(&p.in, &o, &l, &ao, checkpoint_interval,
                                             current_stream(p.in.volume.device)));
          },
          py::arg("inputs"), py::arg("outputs"), py::arg("loss"), py::arg("target"), py::arg("adj_outputs"),
          py::arg("loss_out") = py::none(), py::arg("scale") = 0.0f, py::arg("checkpoint_interval") = 0);
    m.def("render_forward_gradients",
          [](const RendererInputs& inputs, const ForwardDifferencesSettings& fds, RendererOutputs& outputs,
             py::object gradients_out) {
              PackedInputs p = pack_inputs(inputs);
              dvr_outputs o = pack_outputs(outputs);
              dvr_forward_differences f = pack_differences(fds);
              dvr_tensor g = to_tensor(gradients_out);
              check(dvr_render_forward_gradients(&p.in, &f, &o, &g, current_stream(p.in.volume.device)));
          },
          py::arg("inputs"), py::arg("differences_settings"), py::arg("outputs"), py::arg("gradients_out"));
    m.def("forward_variables_to_gradients",
          [](py::object forward_variables, py::object adj_color, const ForwardDifferencesSettings& fds,
             AdjointOutputs& adj) {
              dvr_tensor fv = to_tensor(forward_variables), ac = to_tensor(adj_color);
              dvr_forward_differences f = pack_differences(fds);
              dvr_adjoint_outputs ao = pack_adjoint(adj);
              check(dvr_forward_variables_to_gradients(&fv, &ac, &f, &ao, current_stream(fv.device)));
          },
          py::arg("forward_variables"), py::arg("adj_color"), py::arg("differences_settings"), py::arg("adj_outputs"));
    m.def("compare_to_image",
          [](py::object color, py::object gradients, py::object reference, py::object difference_out,
             py::object gradients_out, bool reduce) {
              dvr_tensor c = to_tensor(color), g = to_tensor(gradients), r = to_tensor(reference);
              dvr_tensor d = to_tensor(difference_out), go = to_tensor(gradients_out);
              check(dvr_compare_to_image(&c, &g, &r, &d, &go, reduce ? 1 : 0, current_stream(c.device)));
          },
          py::arg("color_input"), py::arg("gradients_input"), py::arg("color_reference"), py::arg("difference_out"),
          py::arg("gradients_out"), py::arg("reduce"));
    // TF utilities, importance sampling and elementwise losses on device tensors
    m.def("preshade_volume", [](py::object density, py::object tf, int tf_mode, py::object color) {
        dvr_tensor d = to_tensor(density), t = to_tensor(tf), c = to_tensor(color);
        check(dvr_preshade_volume(&d, &t, tf_mode, &c, current_stream(d.device)));
    });
    m.def("preshade_volume_adjoint",
          [](py::object density, py::object tf, int tf_mode, py::object adj_color, py::object adj_density) {
              dvr_tensor d = to_tensor(density), t = to_tensor(tf), a = to_tensor(adj_color);
              dvr_tensor o = to_tensor(adj_density);
              check(dvr_preshade_volume_adjoint(&d, &t, tf_mode, &a, &o, current_stream(d.device)));
          });
    m.def("find_best_fit", [](py::object color_volume, py::object tf, int tf_mode, int num_samples,
                              float opacity_weighting, py::object previous, float neighbor_weighting, uint64_t seed,
                              py::object density) {
        dvr_tensor c = to_tensor(color_volume), t = to_tensor(tf), p = to_tensor(previous), d = to_tensor(density);
        check(dvr_find_best_fit(&c, &t, tf_mode, num_samples, opacity_weighting, &p, neighbor_weighting, seed, &d,
                                current_stream(c.device)));
    });
    m.def("sample_importance", [](py::object locations, py::object density, py::object tf, int tf_mode,
                                  std::vector<float> weights, int seed, py::object maxima, py::object out) {
        if (weights.size() != 4) throw std::invalid_argument("4 weights expected");
        dvr_tensor l = to_tensor(locations), d = to_tensor(density), t = to_tensor(tf);
        dvr_tensor mx = to_tensor(maxima), o = to_tensor(out);
        check(dvr_sample_importance(&l, &d, &t, tf_mode, weights.data(), seed, &mx, &o, current_stream(d.device)));
    });
    m.def("mul_log", [](py::object t, py::object out) {
        dvr_tensor a = to_tensor(t), o = to_tensor(out);
        check(dvr_mul_log(&a, &o, current_stream(a.device)));
    });
    m.def("mul_log_backward", [](py::object t, py::object grad, py::object grad_t) {
        dvr_tensor a = to_tensor(t), g = to_tensor(grad), o = to_tensor(grad_t);
        check(dvr_mul_log_backward(&a, &g, &o, current_stream(a.device)));
    });
    m.def("log_loss", [](bool l1, py::object x, py::object y, py::object out) {
        dvr_tensor a = to_tensor(x), b = to_tensor(y), o = to_tensor(out);
        check(l1 ? dvr_log_l1(&a, &b, &o, current_stream(a.device)) : dvr_log_mse(&a, &b, &o, current_stream(a.device)));
    });
    m.def("log_loss_backward", [](bool l1, py::object x, py::object y, py::object grad, py::object gx, py::object gy) {
        dvr_tensor a = to_tensor(x), b = to_tensor(y), g = to_tensor(grad), ox = to_tensor(gx), oy = to_tensor(gy);
        void* s = current_stream(a.device);
        check(l1 ? dvr_log_l1_backward(&a, &b, &g, &ox, &oy, s) : dvr_log_mse_backward(&a, &b, &g, &ox, &oy, s));
    });
    // volume files, synthetic volumes, histogram and TF builders (host buffers by address)
    m.def("cvol_read_header", [](const std::string& path) {
        dvr_cvol_header h;
        check(dvr_cvol_read_header(path.c_str(), &h));
        return py::make_tuple(py::make_tuple(h.size[0], h.size[1], h.size[2]),
                              py::make_tuple(h.voxel_size[0], h.voxel_size[1], h.voxel_size[2]), h.type,
                              h.compressed != 0);
    });
    m.def("cvol_read", [](const std::string& path, uintptr_t data, int64_t count) {
        check(dvr_cvol_read(path.c_str(), reinterpret_cast<float*>(data), count));
    });
    m.def("cvol_write", [](const std::string& path, uintptr_t data, std::vector<int64_t> size, std::vector<double> world,
                           int compression) {
        if (size.size() != 3 || world.size() != 3) throw std::invalid_argument("size and world size need 3 entries");
        check(dvr_cvol_write(path.c_str(), reinterpret_cast<const float*>(data), size.data(), world.data(), compression));
    });
    m.def("volume_implicit", [](int equation, int resolution, std::vector<std::string> names, std::vector<float> values,
                                uintptr_t data) {
        if (names.size() != values.size()) throw std::invalid_argument("names and values differ in length");
        std::vector<const char*> cn;
        for (auto& n : names) cn.push_back(n.c_str());
        check(dvr_volume_implicit(equation, resolution, (int)cn.size(), cn.data(), values.data(),
                                  reinterpret_cast<float*>(data)));
    });
    m.def("volume_histogram", [](uintptr_t data, std::vector<int64_t> size, std::vector<int64_t> strides) {
        if (size.size() != 3 || strides.size() != 3) throw std::invalid_argument("size and strides need 3 entries");
        std::vector<float> bins(DVR_HISTOGRAM_BINS);
        float mn = 0, mx = 0;
        uint32_t nnz = 0;
        check(dvr_volume_histogram(reinterpret_cast<const float*>(data), size.data(), strides.data(), bins.data(), &mn,
                                   &mx, &nnz));
        return py::make_tuple(bins, mn, mx, nnz);
    });
    m.def("rgb_to_lab", [](float r, float g, float b) {
        const float in[3] = {r, g, b};
        float out[3];
        dvr_rgb_to_lab(in, out);
        return py::make_tuple(out[0], out[1], out[2]);
    });
    m.def("lab_to_rgb", [](float l, float a, float b) {
        const float in[3] = {l, a, b};
        float out[3];
        dvr_lab_to_rgb(in, out);
        return py::make_tuple(out[0], out[1], out[2]);
    });
    m.def("tf_assemble_from_settings", [](std::vector<float> color_lab, std::vector<float> color_pos,
                                          std::vector<float> opacity, std::vector<float> opacity_pos, float min_density,
                                          float max_density, float opacity_scaling, bool purge) {
        if (color_lab.size() != 3 * color_pos.size()) throw std::invalid_argument("colour values need 3 entries each");
        if (opacity.size() != opacity_pos.size()) throw std::invalid_argument("opacity values and positions differ");
        int32_t count = 0;
        check(dvr_tf_assemble_from_settings(color_lab.data(), color_pos.data(), (int32_t)color_pos.size(),
                                            opacity.data(), opacity_pos.data(), (int32_t)opacity.size(), min_density,
                                            max_density, opacity_scaling, purge ? 1 : 0, nullptr, 0, &count));
        std::vector<dvr_tf_point> pts((size_t)count);
        check(dvr_tf_assemble_from_settings(color_lab.data(), color_pos.data(), (int32_t)color_pos.size(),
                                            opacity.data(), opacity_pos.data(), (int32_t)opacity.size(), min_density,
                                            max_density, opacity_scaling, purge ? 1 : 0, pts.data(), count, &count));
        py::list out;
        for (auto& p : pts) out.append(py::make_tuple(p.pos, py::make_tuple(p.val[0], p.val[1], p.val[2], p.val[3])));
        return out;
    });
    m.def("tf_texture_from_points", [](std::vector<float> pos, std::vector<float> val, int resolution, uintptr_t out) {
        if (val.size() != 4 * pos.size()) throw std::invalid_argument("values need 4 entries per point");
        std::vector<dvr_tf_point> pts(pos.size());
        for (size_t i = 0; i < pos.size(); ++i) {
            pts[i].pos = pos[i];
            for (int c = 0; c < 4; ++c) pts[i].val[c] = val[4 * i + c];
        }
        check(dvr_tf_texture_from_points(pts.data(), (int32_t)pts.size(), resolution, reinterpret_cast<float*>(out)));
    });
    m.def("checkpoint_buffer_elements",
          [](const RendererInputs& inputs, int K) {
              PackedInputs p = pack_inputs(inputs);
              const int64_t n = dvr_checkpoint_buffer_elements(&p.in, K);
              if (n < 0) throw std::runtime_error(dvr_last_error());
              return n;
          },
          py::arg("inputs"), py::arg("checkpoint_interval"));
    m.def("density_buffer_elements",
          [](const RendererInputs& inputs) {
              PackedInputs p = pack_inputs(inputs);
              const int64_t n = dvr_density_buffer_elements(&p.in);
              if (n < 0) throw std::runtime_error(dvr_last_error());
              return n;
          },
          py::arg("inputs"));
    m.def("generate_rays",
          [](py::object viewport, float fovY, int W, int H, py::object ray_start, py::object ray_dir) {
              dvr_tensor v = to_tensor(viewport), s = to_tensor(ray_start), d = to_tensor(ray_dir);
              check(dvr_generate_rays(&v, fovY, W, H, &s, &d, current_stream(v.device)));
          });
    m.def("generate_rays_adjoint",
          [](py::object viewport, float fovY, int W, int H, py::object adj_start, py::object adj_dir, py::object out) {
              dvr_tensor v = to_tensor(viewport), s = to_tensor(adj_start), d = to_tensor(adj_dir), o = to_tensor(out);
              check(dvr_generate_rays_adjoint(&v, fovY, W, H, &s, &d, &o, current_stream(v.device)));
          });

    py::class_<BatchPlanHandle>(m, "_BatchPlan")
        .def(py::init(&make_plan), py::arg("volumes"), py::arg("tfs"), py::arg("views"), py::arg("filter"),
             py::arg("tf_mode"), py::arg("blend"), py::arg("early_out"), py::arg("loss"), py::arg("loss_out"),
             py::arg("forward_only"), py::arg("device"), py::arg("checkpoint_interval"), py::arg("wave_width"),
             py::arg("grid"), py::arg("reinterpolate") = false)
        .def("run", [](BatchPlanHandle& h, int flags, uintptr_t stream) {
            check(dvr_batch_plan_run(h.plan, flags, reinterpret_cast<void*>(stream)));
        })
        .def("zero_gradients", [](BatchPlanHandle& h, uintptr_t stream) {
            check(dvr_batch_plan_zero_gradients(h.plan, reinterpret_cast<void*>(stream)));
        })
        .def("time_kernel", [](BatchPlanHandle& h, int flags, uintptr_t stream) {
            float ms = 0.0f;
            check(dvr_batch_plan_time_kernel(h.plan, flags, reinterpret_cast<void*>(stream), &ms));
            return ms;
        })
        .def("freeze", [](BatchPlanHandle& h) { check(dvr_batch_plan_freeze(h.plan)); })
        .def("options", &plan_options)
        .def("set_options", [](BatchPlanHandle& h, int K, int ww, int grid, bool reinterpolate) {
            dvr_batch_options o{K, ww, grid, reinterpolate ? 1 : 0};
            check(dvr_batch_plan_set_options(h.plan, &o));
        }, py::arg("checkpoint_interval"), py::arg("wave_width"), py::arg("grid"), py::arg("reinterpolate") = false)
        .def("autotune", [](BatchPlanHandle& h, int iterations, uintptr_t stream) {
            dvr_batch_options o;
            check(dvr_batch_plan_autotune(h.plan, iterations, &o, reinterpret_cast<void*>(stream)));
            return py::make_tuple(o.checkpoint_interval, o.wave_width, o.grid, o.reinterpolate != 0);
        })
        .def("workspace_bytes", [](const BatchPlanHandle& h) { return dvr_batch_plan_workspace_bytes(h.plan); })
        .def("chunk_slots", [](const BatchPlanHandle& h) { return dvr_batch_plan_chunk_slots(h.plan); });

    py::class_<AdamHandle>(m, "_AdamGroup")
        .def(py::init([](py::list blocks, int device) {
                 std::vector<dvr_adam_block> bs;
                 for (auto b : blocks) {
                     py::tuple t = py::reinterpret_borrow<py::tuple>(b);
                     if (t.size() != 9)
                         throw std::invalid_argument("block: (param, grad, exp_avg, exp_avg_sq, rows, cols, stride, "
                                                     "clamp_min, clamp_max)");
                     dvr_adam_block a;
                     a.param = ptr_of<float>(t[0]); a.grad = ptr_of<const float>(t[1]);
                     a.exp_avg = ptr_of<float>(t[2]); a.exp_avg_sq = ptr_of<float>(t[3]);
                     a.rows = t[4].cast<int64_t>(); a.cols = t[5].cast<int64_t>(); a.stride = t[6].cast<int64_t>();
                     a.clamp_min = t[7].cast<float>(); a.clamp_max = t[8].cast<float>();
                     bs.push_back(a);
                 }
                 std::unique_ptr<AdamHandle> h(new AdamHandle());
                 check(dvr_adam_group_create(bs.data(), (int)bs.size(), device, &h->group));
                 return h;
             }),
             py::arg("blocks"), py::arg("device"))
        .def("step", [](AdamHandle& h, double lr, double b1, double b2, double eps, uintptr_t stream) {
            check(dvr_adam_group_step(h.group, lr, b1, b2, eps, reinterpret_cast<void*>(stream)));
        })
        .def("set_step", [](AdamHandle& h, int64_t step) { check(dvr_adam_group_set_step(h.group, step)); })
        .def("get_step", [](AdamHandle& h) {
            int64_t s = 0;
            check(dvr_adam_group_get_step(h.group, &s));
            return s;
        });
}
