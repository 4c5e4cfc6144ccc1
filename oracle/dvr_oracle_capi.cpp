// C entry points of the CPU oracle (TEST INFRASTRUCTURE ONLY).
// Loaded with ctypes by tests/, by __graft_entry__.smoke() and by the
// cpu_baseline leg of bench.py; never by the product library.
//
// Every function exists in an FP64 (suffix _f64) and an FP32 (suffix _f32)
// flavour; pointer arguments are of the matching scalar type.
#include "dvr_oracle.hpp"

#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {

// Mirrors oracle::Inputs with type-erased pointers and double scalars.
typedef struct {
    int B, W, H;
    int filter, camera, tf_mode, blend;
    const void* volume; int vol_b, X, Y, Z;
    const void* box_min; int box_min_b;
    const void* box_size; int box_size_b;
    const void* ray_start; const void* ray_dir; int ray_b;
    const void* cam_matrix; int cam_b;
    double fov_y;
    const void* step; int step_b, step_h, step_w;
    const void* tf; int tf_b, R, C;
    double early_out;
} oracle_inputs;

typedef struct {
    void* adj_volume; int adj_volume_b;
    void* adj_tf; int adj_tf_b;
    void* adj_ray_start; void* adj_ray_dir; int adj_cam_b;
    void* adj_step; int adj_step_b, adj_step_h, adj_step_w;
} oracle_adjoint_outputs;

}  // extern "C"

namespace {
template <class T> oracle::Inputs<T> convert(const oracle_inputs* s)
{
    oracle::Inputs<T> in;
    in.B = s->B; in.W = s->W; in.H = s->H;
    in.filter = s->filter; in.camera = s->camera; in.tf_mode = s->tf_mode; in.blend = s->blend;
    in.volume = (const T*)s->volume; in.vol_b = s->vol_b; in.X = s->X; in.Y = s->Y; in.Z = s->Z;
    in.box_min = (const T*)s->box_min; in.box_min_b = s->box_min_b;
    in.box_size = (const T*)s->box_size; in.box_size_b = s->box_size_b;
    in.ray_start = (const T*)s->ray_start; in.ray_dir = (const T*)s->ray_dir; in.ray_b = s->ray_b;
    in.cam_matrix = (const T*)s->cam_matrix; in.cam_b = s->cam_b;
    in.fov_y = (T)s->fov_y;
    in.step = (const T*)s->step; in.step_b = s->step_b; in.step_h = s->step_h; in.step_w = s->step_w;
    in.tf = (const T*)s->tf; in.tf_b = s->tf_b; in.R = s->R; in.C = s->C;
    in.early_out = (T)s->early_out;
    return in;
}
template <class T> oracle::AdjointOutputs<T> convert(const oracle_adjoint_outputs* s)
{
    oracle::AdjointOutputs<T> o;
    o.adj_volume = (T*)s->adj_volume; o.adj_volume_b = s->adj_volume_b;
    o.adj_tf = (T*)s->adj_tf; o.adj_tf_b = s->adj_tf_b;
    o.adj_ray_start = (T*)s->adj_ray_start; o.adj_ray_dir = (T*)s->adj_ray_dir; o.adj_cam_b = s->adj_cam_b;
    o.adj_step = (T*)s->adj_step; o.adj_step_b = s->adj_step_b; o.adj_step_h = s->adj_step_h; o.adj_step_w = s->adj_step_w;
    return o;
}

// Component harness: a one-pixel Inputs carrying only a TF table.
template <class T> oracle::Inputs<T> tf_only(int mode, const T* tf, int R, int C)
{
    oracle::Inputs<T> in{};
    in.tf_mode = mode; in.tf = tf; in.tf_b = 1; in.R = R; in.C = C;
    return in;
}
}  // namespace

#define ORACLE_DEFINE(T, SUF)                                                                                    \
    extern "C" void oracle_render_forward_##SUF(const oracle_inputs* in, T* color, int* term,                  \
                                                  const int* frozen_term, const T* frozen_t0, T* t0_out)         \
    {                                                                                                          \
        oracle::render_forward<T>(convert<T>(in), color, term, frozen_term, frozen_t0, t0_out);                \
    }                                                                                                          \
    /* camera rays of every pixel (a2): start, dir B*H*W*3 */                                              \
    extern "C" void oracle_camera_rays_##SUF(const oracle_inputs* in, T* start, T* dir)                       \
    {                                                                                                          \
        const oracle::Inputs<T> c = convert<T>(in);                                                            \
        for (int b = 0; b < c.B; ++b)                                                                          \
            for (int y = 0; y < c.H; ++y)                                                                      \
                for (int x = 0; x < c.W; ++x) {                                                                \
                    oracle::V3<T> s, d;                                                                        \
                    oracle::camera_rays<T>(c, x, y, b, s, d);                                                  \
                    const size_t o = (((size_t)b * c.H + y) * c.W + x) * 3;                                    \
                    start[o] = s.x; start[o + 1] = s.y; start[o + 2] = s.z;                                    \
                    dir[o] = d.x; dir[o + 1] = d.y; dir[o + 2] = d.z;                                          \
                }                                                                                              \
    }                                                                                                          \
    extern "C" void oracle_kink_census_##SUF(const oracle_inputs* in, const int* term, T eps, T eps_pos,        \
                                             T eps_face, int* out)                                             \
    {                                                                                                          \
        oracle::kink_census<T>(convert<T>(in), term, eps, eps_pos, eps_face, out);                             \
    }                                                                                                          \
    /* forward-mode derivatives: grad is B*H*W*D*4; d_start/d_dir hold 3 indices; d_tf may be null */          \
    extern "C" void oracle_render_forward_gradients_##SUF(const oracle_inputs* in, int D, int d_step,            \
                                                            const int* d_start, const int* d_dir,                \
                                                            const int* d_tf, int d_tf_b, T* color, int* term,    \
                                                            T* grad, const int* frozen_term, const T* frozen_t0)  \
    {                                                                                                          \
        oracle::FwdDiffSettings fd{D, d_step, {d_start[0], d_start[1], d_start[2]},                            \
                                   {d_dir[0], d_dir[1], d_dir[2]}, d_tf, d_tf_b};                              \
        oracle::render_forward_gradients<T>(convert<T>(in), fd, color, term, grad, frozen_term, frozen_t0);    \
    }                                                                                                          \
    extern "C" void oracle_render_adjoint_##SUF(const oracle_inputs* in, const T* color, const int* term,       \
                                                  const T* adj_color, const oracle_adjoint_outputs* out,          \
                                                  int method, const T* frozen_t0)                                 \
    {                                                                                                          \
        oracle::render_adjoint<T>(convert<T>(in), color, term, adj_color, convert<T>(out), method, frozen_t0);  \
    }                                                                                                          \
    extern "C" void oracle_blend_##SUF(int mode, const T* acc, const T* cur, T step, T* out)                    \
    {                                                                                                          \
        oracle::V4<T> r = oracle::blend<T>(mode, oracle::v4(acc[0], acc[1], acc[2], acc[3]),                   \
                                           oracle::v4(cur[0], cur[1], cur[2], cur[3]), step);                  \
        out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;                                                \
    }                                                                                                          \
    /* inversion adjoint of one blend step: acc_out -> acc_in, adjoints of (acc_in, cur, step) */               \
    extern "C" void oracle_blend_adjoint_##SUF(int mode, const T* acc_out, const T* cur, T step,                \
                                               const T* adj_out, T* acc_in, T* adj_acc_in, T* adj_cur,           \
                                               T* adj_step)                                                      \
    {                                                                                                          \
        const oracle::V4<T> c = oracle::v4(cur[0], cur[1], cur[2], cur[3]);                                    \
        const oracle::V4<T> ai = oracle::blend_invert<T>(mode, oracle::v4(acc_out[0], acc_out[1], acc_out[2],  \
                                                                          acc_out[3]), c, step);             \
        oracle::V4<T> adj = oracle::v4(adj_out[0], adj_out[1], adj_out[2], adj_out[3]);                        \
        oracle::V4<T> ac;                                                                                      \
        T as;                                                                                                  \
        oracle::blend_adjoint_given<T>(mode, ai.w, c, step, adj, ac, as);                                      \
        acc_in[0] = ai.x; acc_in[1] = ai.y; acc_in[2] = ai.z; acc_in[3] = ai.w;                                \
        adj_acc_in[0] = adj.x; adj_acc_in[1] = adj.y; adj_acc_in[2] = adj.z; adj_acc_in[3] = adj.w;            \
        adj_cur[0] = ac.x; adj_cur[1] = ac.y; adj_cur[2] = ac.z; adj_cur[3] = ac.w;                            \
        *adj_step = as;                                                                                        \
    }                                                                                                          \
    extern "C" void oracle_tf_eval_##SUF(int mode, const T* tf, int R, int C, T density, T* out)                \
    {                                                                                                          \
        oracle::V4<T> r = oracle::tf_eval<T>(tf_only<T>(mode, tf, R, C), 0, density);                          \
        out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;                                                \
    }                                                                                                          \
    /* adj_tf (R*C) is accumulated into; may be null */                                                        \
    extern "C" T oracle_tf_adjoint_##SUF(int mode, const T* tf, int R, int C, T density, const T* adj,          \
                                          T* adj_tf)                                                            \
    {                                                                                                          \
        return oracle::tf_adjoint<T>(tf_only<T>(mode, tf, R, C), 0, density,                                   \
                                     oracle::v4(adj[0], adj[1], adj[2], adj[3]), adj_tf);                      \
    }                                                                                                          \
    /* trilinear (filter=1) or nearest (filter=0) fetch from a dense X*Y*Z volume */                           \
    extern "C" T oracle_fetch_##SUF(int filter, const T* vol, int X, int Y, int Z, const T* p)                  \
    {                                                                                                          \
        oracle::Inputs<T> in{};                                                                                \
        in.filter = filter; in.volume = vol; in.vol_b = 1; in.X = X; in.Y = Y; in.Z = Z;                       \
        return oracle::fetch_density<T>(in, 0, oracle::v3(p[0], p[1], p[2]));                                  \
    }                                                                                                          \
    /* trilinear adjoint: accumulates into adj_vol (may be null), writes adj_p[3] */                          \
    extern "C" void oracle_trilinear_adjoint_##SUF(const T* vol, int X, int Y, int Z, const T* p, T adj_d,     \
                                                   T* adj_vol, T* adj_p)                                       \
    {                                                                                                          \
        oracle::Inputs<T> in{};                                                                                \
        in.volume = vol; in.vol_b = 1; in.X = X; in.Y = Y; in.Z = Z;                                           \
        oracle::TriCell<T> c;                                                                                  \
        oracle::trilinear_cell<T>(in, 0, oracle::v3(p[0], p[1], p[2]), c);                                     \
        T w[8];                                                                                                \
        oracle::V3<T> ap = oracle::trilinear_adjoint<T>(c, adj_d, w);                                          \
        if (adj_vol)                                                                                           \
            for (int k = 0; k < 8; ++k) {                                                                      \
                const int ax = (k >> 2) & 1 ? c.hx : c.lx, ay = (k >> 1) & 1 ? c.hy : c.ly,                     \
                          az = k & 1 ? c.hz : c.lz;                                                            \
                adj_vol[((size_t)ax * Y + ay) * Z + az] += w[k];                                               \
            }                                                                                                  \
        adj_p[0] = ap.x; adj_p[1] = ap.y; adj_p[2] = ap.z;                                                     \
    }                                                                                                          \
    extern "C" void oracle_adj_cross_##SUF(const T* a, const T* b, const T* adj_c, T* adj_a, T* adj_b)         \
    {                                                                                                          \
        oracle::V3<T> aa = oracle::v3(adj_a[0], adj_a[1], adj_a[2]), ab = oracle::v3(adj_b[0], adj_b[1], adj_b[2]); \
        oracle::adj_cross<T>(oracle::v3(a[0], a[1], a[2]), oracle::v3(b[0], b[1], b[2]),                       \
                             oracle::v3(adj_c[0], adj_c[1], adj_c[2]), aa, ab);                                \
        adj_a[0] = aa.x; adj_a[1] = aa.y; adj_a[2] = aa.z; adj_b[0] = ab.x; adj_b[1] = ab.y; adj_b[2] = ab.z;  \
    }                                                                                                          \
    extern "C" void oracle_adj_normalize_##SUF(const T* a, const T* adj_out, T* adj_a)                         \
    {                                                                                                          \
        oracle::V3<T> r = oracle::adj_normalize<T>(oracle::v3(a[0], a[1], a[2]),                               \
                                                   oracle::v3(adj_out[0], adj_out[1], adj_out[2]));            \
        adj_a[0] = r.x; adj_a[1] = r.y; adj_a[2] = r.z;                                                        \
    }                                                                                                          \
    extern "C" void oracle_generate_rays_##SUF(const T* viewport, int B, int H, int W, T fov_y, T* start,     \
                                               T* dir)                                                         \
    {                                                                                                          \
        oracle::generate_rays<T>(viewport, B, H, W, fov_y, start, dir);                                        \
    }                                                                                                          \
    extern "C" void oracle_generate_rays_adjoint_##SUF(const T* viewport, int B, int H, int W, T fov_y,        \
                                                       const T* adj_start, const T* adj_dir, T* adj_viewport)  \
    {                                                                                                          \
        oracle::generate_rays_adjoint<T>(viewport, B, H, W, fov_y, adj_start, adj_dir, adj_viewport);          \
    }

#define ORACLE_DEFINE_AUX(T, SUF)                                                                                \
    extern "C" void oracle_preshade_##SUF(int mode, const T* tf, int R, int C, const T* dens, int64_t n, T* color) \
    {                                                                                                          \
        oracle::preshade<T>(tf_only<T>(mode, tf, R, C), dens, n, color);                                       \
    }                                                                                                          \
    extern "C" void oracle_preshade_adjoint_##SUF(int mode, const T* tf, int R, int C, const T* dens, int64_t n,  \
                                                  const T* adj_color, T* adj_dens)                              \
    {                                                                                                          \
        oracle::preshade_adjoint<T>(tf_only<T>(mode, tf, R, C), dens, n, adj_color, adj_dens);                 \
    }                                                                                                          \
    extern "C" void oracle_best_fit_##SUF(int mode, const T* tf, int R, int C, const T* color, int X, int Y,      \
                                          int Z, int num_samples, T opacity_w, const T* prev, T neigh_w,         \
                                          uint64_t seed, T* out, T* best_cost)                                   \
    {                                                                                                          \
        oracle::best_fit<T>(tf_only<T>(mode, tf, R, C), color, X, Y, Z, num_samples, opacity_w, prev, neigh_w,   \
                            seed, out, best_cost);                                                             \
    }                                                                                                          \
    /* nb: 6 neighbour densities (x-, x+, y-, y+, z-, z+) or null; has: 6 flags */                             \
    extern "C" T oracle_best_fit_cost_##SUF(int mode, const T* tf, int R, int C, const T* target, T d,          \
                                            T opacity_w, const T* nb, const int* has, T neigh_w)                \
    {                                                                                                          \
        bool h[6] = {false, false, false, false, false, false};                                               \
        if (has)                                                                                               \
            for (int j = 0; j < 6; ++j) h[j] = has[j] != 0;                                                    \
        return oracle::best_fit_cost<T>(tf_only<T>(mode, tf, R, C), target, d, opacity_w, nb, h, neigh_w);     \
    }                                                                                                          \
    extern "C" void oracle_importance_maxima_##SUF(int mode, const T* tf, int R, int C, const T* dens, int X,   \
                                                   int Y, int Z, T* out)                                        \
    {                                                                                                          \
        oracle::importance_maxima<T>(tf_only<T>(mode, tf, R, C), dens, X, Y, Z, out);                          \
    }                                                                                                          \
    extern "C" void oracle_sample_importance_##SUF(int mode, const T* tf, int R, int C, const T* dens, int X,   \
                                                   int Y, int Z, const T* loc, int64_t N, const float* w,       \
                                                   const T* maxima, uint32_t seed, uint8_t* out, T* margin)     \
    {                                                                                                          \
        oracle::sample_importance<T>(tf_only<T>(mode, tf, R, C), dens, X, Y, Z, loc, N, w, maxima, seed, out,  \
                                     margin);                                                                  \
    }

ORACLE_DEFINE(double, f64)
ORACLE_DEFINE(float, f32)
ORACLE_DEFINE_AUX(double, f64)
ORACLE_DEFINE_AUX(float, f32)

extern "C" uint64_t oracle_splitmix64(uint64_t x) { return oracle::splitmix64(x); }
extern "C" uint64_t oracle_lcg48_next(uint64_t s) { return oracle::lcg48_next(s); }
extern "C" uint64_t oracle_best_fit_stream(uint64_t seed, uint64_t voxel) { return oracle::best_fit_stream(seed, voxel); }

extern "C" void oracle_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

extern "C" int oracle_get_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
