// CPU oracle for differentiable direct volume rendering (forward ray march +
// adjoint).  TEST INFRASTRUCTURE ONLY: this code is the checker that the HIP
// kernels are compared against.  It is never linked into the product library.
//
// It is an independent, scalar C++ restatement of the reference algorithm,
// templated on the scalar type: T=double is the FP64 oracle, T=float is the
// FP32 twin that follows the reference's single-precision operation order and
// doubles as the OpenMP CPU baseline (the reference's renderer_cpu.cpp path,
// which cannot be built here).
//
// Reference files followed (paths relative to the DiffDVR repository):
//   renderer/renderer_kernels.cuh:22-127   forward march (DvrKernelForwardImpl/Host)
//   renderer/renderer_kernels.cuh:315-497  adjoint march (DvrKernelAdjointImpl/Host)
//   renderer/renderer_camera.cuh:12-61     camera modes
//   renderer/renderer_utils.cuh:53-74      matmul_proj, intersectionRayAABB
//   renderer/renderer_interpolation.cuh:33-389  nearest / trilinear (+adjoint)
//   renderer/renderer_tf.cuh:59-648        identity / texture / linear / gaussian TF (+adjoint)
//   renderer/renderer_blending.cuh:36-165  Beer-Lambert / alpha blending (+adjoint by inversion)
//   renderer/renderer_adjoint.cuh:28-65    adjNormalize, adjCross
//   renderer/camera_impl.cu:17-174,302-357 generateRays forward / adjoint
//   renderer/helper_math.cuh:86-104,1898-1909  rmin/rmax (ternary), clamp (fmin/fmax)
//
// Deliberate deviations (documented in the design notes):
//   * the step-size adjoint is accumulated over all samples; the reference
//     overwrites its running sum each iteration (renderer_kernels.cuh:393-395
//     with renderer_blending.cuh:93), which leaves only sample 0's term.
//   * the Gaussian-TF sigma adjoint keeps the scalar type; the reference
//     rounds one factor through `float` (renderer_tf.cuh:591).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

namespace oracle {

enum FilterMode { FilterNearest = 0, FilterTrilinear = 1, FilterPreshaded = 3 };
enum CameraMode { CameraRayStartDir = 0, CameraInverseViewMatrix = 1, CameraReferenceFrame = 2 };
enum TFMode { TFIdentity = 0, TFTexture = 1, TFLinear = 2, TFGaussian = 3, TFPreshaded = 5, TFGaussianLog = 6 };
enum BlendMode { BlendBeerLambert = 0, BlendAlpha = 1 };
enum AdjointMethod { AdjointInversion = 0, AdjointExactReplay = 1 };

template <class T> struct V3 { T x, y, z; };
template <class T> struct V4 { T x, y, z, w; };

template <class T> inline V3<T> v3(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> inline V4<T> v4(T x, T y, T z, T w) { return V4<T>{x, y, z, w}; }
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator*(V3<T> a, V3<T> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class T> inline V3<T> operator/(V3<T> a, V3<T> b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
template <class T> inline V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline T dot(V4<T> a, V4<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
template <class T> inline V3<T> cross(V3<T> a, V3<T> b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// helper_math.cuh:86-104 -- strict ternary min/max (not fmin/fmax)
template <class T> inline T tmin_(T a, T b) { return a < b ? a : b; }
template <class T> inline T tmax_(T a, T b) { return a > b ? a : b; }
// helper_math.cuh:1898-1909 -- clamp(f,a,b) = fmax(a, fmin(f, b))
template <class T> inline T fclamp(T f, T a, T b) { return std::fmax(a, std::fmin(f, b)); }
inline int iclamp(int f, int a, int b) { return std::max(a, std::min(f, b)); }
template <class T> inline T rsqrt_(T x) { return T(1) / std::sqrt(x); }
template <class T> inline V3<T> normalize(V3<T> v)
{
    T inv = rsqrt_(dot(v, v));
    return v * inv;
}
template <class T> inline T length(V3<T> v) { return std::sqrt(dot(v, v)); }

// Atomic scatter: the reference uses `#pragma omp atomic` on the CPU
// (renderer_adjoint.cuh:17-26).
template <class T> inline void atomic_add(T* p, T v)
{
#pragma omp atomic
    *p += v;
}

// ---------------------------------------------------------------------------
// Inputs / outputs.  Every batched input carries its own batch count nb
// (1 = broadcast over the render batch, renderer_tensor.cuh:140-144).
// All arrays are dense and C-contiguous.
// ---------------------------------------------------------------------------
template <class T> struct Inputs {
    int B, W, H;
    int filter, camera, tf_mode, blend;
    const T* volume; int vol_b, X, Y, Z;        // vol_b x X x Y x Z
    const T* box_min; int box_min_b;            // box_min_b x 3
    const T* box_size; int box_size_b;          // box_size_b x 3
    const T* ray_start; const T* ray_dir; int ray_b;  // ray_b x H x W x 3
    const T* cam_matrix; int cam_b;             // cam_b x 4 x 4  or  cam_b x 3 x 3
    T fov_y;
    const T* step; int step_b, step_h, step_w;  // step_b x step_h x step_w (each 1 or full)
    const T* tf; int tf_b, R, C;                // tf_b x R x C
    T early_out;
};

template <class T> struct AdjointOutputs {
    T* adj_volume; int adj_volume_b;            // null if not requested
    T* adj_tf; int adj_tf_b;
    T* adj_ray_start; T* adj_ray_dir; int adj_cam_b;
    T* adj_step; int adj_step_b, adj_step_h, adj_step_w;
};

inline int bidx(int nb, int b) { return nb == 1 ? 0 : b; }

// ---------------------------------------------------------------------------
// a2: camera evaluation (renderer_camera.cuh:12-61; renderer_utils.cuh:43-58)
// ---------------------------------------------------------------------------
template <class T>
inline V3<T> matmul_proj(const T* m /*4x4 row-major*/, T x, T y, T z)
{
    T s[4] = {x, y, z, T(1)};
    T w[4];
    for (int r = 0; r < 4; ++r)
        w[r] = m[4 * r + 0] * s[0] + m[4 * r + 1] * s[1] + m[4 * r + 2] * s[2] + m[4 * r + 3] * s[3];
    return v3<T>(w[0] / w[3], w[1] / w[3], w[2] / w[3]);
}

template <class T>
inline void camera_rays(const Inputs<T>& in, int x, int y, int b, V3<T>& start, V3<T>& dir)
{
    if (in.camera == CameraRayStartDir) {
        const int cb = bidx(in.ray_b, b);
        const size_t o = (((size_t)cb * in.H + y) * in.W + x) * 3;
        start = v3(in.ray_start[o], in.ray_start[o + 1], in.ray_start[o + 2]);
        dir = v3(in.ray_dir[o], in.ray_dir[o + 1], in.ray_dir[o + 2]);
    } else if (in.camera == CameraInverseViewMatrix) {
        const T* m = in.cam_matrix + (size_t)bidx(in.cam_b, b) * 16;
        const T fx = T(2) * (T(x) + T(0.5)) / T(in.W) - T(1);
        const T fy = T(2) * (T(y) + T(0.5)) / T(in.H) - T(1);
        start = matmul_proj(m, fx, fy, T(0));
        const V3<T> end = matmul_proj(m, fx, fy, T(0.9));
        dir = normalize(end - start);
    } else {
        const T* m = in.cam_matrix + (size_t)bidx(in.cam_b, b) * 9;
        const T fovX = in.fov_y * T(in.W) / T(in.H);
        const T tanX = std::tan(fovX / T(2));
        const T tanY = std::tan(in.fov_y / T(2));
        const V3<T> eye = v3(m[0], m[1], m[2]);
        const V3<T> right = v3(m[3], m[4], m[5]);
        const V3<T> up = v3(m[6], m[7], m[8]);
        const V3<T> front = cross(up, right);
        const T fx = T(2) * (T(x) + T(0.5)) / T(in.W) - T(1);
        const T fy = T(2) * (T(y) + T(0.5)) / T(in.H) - T(1);
        V3<T> d = front + (fx * tanX) * right + (fy * tanY) * up;
        start = eye;
        dir = normalize(d);
    }
}

// a3: slab test (renderer_utils.cuh:60-74)
template <class T>
inline void ray_aabb(V3<T> start, V3<T> dir, V3<T> bmin, V3<T> bsize, T& t0, T& t1)
{
    const V3<T> inv = v3(T(1) / dir.x, T(1) / dir.y, T(1) / dir.z);
    const T a1 = (bmin.x - start.x) * inv.x, a2 = (bmin.x + bsize.x - start.x) * inv.x;
    const T a3 = (bmin.y - start.y) * inv.y, a4 = (bmin.y + bsize.y - start.y) * inv.y;
    const T a5 = (bmin.z - start.z) * inv.z, a6 = (bmin.z + bsize.z - start.z) * inv.z;
    t0 = tmax_(tmax_(tmin_(a1, a2), tmin_(a3, a4)), tmin_(a5, a6));
    t1 = tmin_(tmin_(tmax_(a1, a2), tmax_(a3, a4)), tmax_(a5, a6));
}

// ---------------------------------------------------------------------------
// a4/a10: volume interpolation (renderer_interpolation.cuh:33-389)
// ---------------------------------------------------------------------------
template <class T>
inline T vol_at(const Inputs<T>& in, int vb, int ix, int iy, int iz)
{
    return in.volume[(((size_t)vb * in.X + ix) * in.Y + iy) * in.Z + iz];
}

// Nearest: clamp to [eps, size-1-eps], round half away from zero (:38-50)
template <class T>
inline void nearest_index(const Inputs<T>& in, V3<T> p, int& ix, int& iy, int& iz)
{
    const T eps = T(1e-7);
    const T cx = fclamp(p.x, eps, T(in.X) - (T(1) + eps));
    const T cy = fclamp(p.y, eps, T(in.Y) - (T(1) + eps));
    const T cz = fclamp(p.z, eps, T(in.Z) - (T(1) + eps));
    ix = (int)std::round(cx);
    iy = (int)std::round(cy);
    iz = (int)std::round(cz);
}

// Trilinear: cell = trunc(p) (int cast), corners clamped, fraction unclamped,
// lerp a+t(b-a) in x, then y, then z (:145-177, :331-357)
template <class T> struct TriCell {
    int ix, iy, iz;       // truncated cell index
    int lx, ly, lz, hx, hy, hz;  // clamped corner indices
    T d[8];               // corner densities, bit order (x<<2)|(y<<1)|z
    T fx, fy, fz;
};
template <class T> inline T lerp_(T a, T b, T t) { return a + t * (b - a); }

template <class T>
inline void trilinear_cell(const Inputs<T>& in, int vb, V3<T> p, TriCell<T>& c)
{
    c.ix = (int)p.x; c.iy = (int)p.y; c.iz = (int)p.z;
    c.lx = iclamp(c.ix, 0, in.X - 1); c.hx = iclamp(c.ix + 1, 0, in.X - 1);
    c.ly = iclamp(c.iy, 0, in.Y - 1); c.hy = iclamp(c.iy + 1, 0, in.Y - 1);
    c.lz = iclamp(c.iz, 0, in.Z - 1); c.hz = iclamp(c.iz + 1, 0, in.Z - 1);
    for (int k = 0; k < 8; ++k) {
        const int ax = (k >> 2) & 1 ? c.hx : c.lx;
        const int ay = (k >> 1) & 1 ? c.hy : c.ly;
        const int az = k & 1 ? c.hz : c.lz;
        c.d[k] = vol_at(in, vb, ax, ay, az);
    }
    c.fx = p.x - T(c.ix); c.fy = p.y - T(c.iy); c.fz = p.z - T(c.iz);
}

template <class T> inline T trilinear_value(const TriCell<T>& c)
{
    const T* d = c.d;
    return lerp_(lerp_(lerp_(d[0], d[4], c.fx), lerp_(d[2], d[6], c.fx), c.fy),
                 lerp_(lerp_(d[1], d[5], c.fx), lerp_(d[3], d[7], c.fx), c.fy), c.fz);
}

template <class T>
inline T fetch_density(const Inputs<T>& in, int b, V3<T> p)
{
    const int vb = bidx(in.vol_b, b);
    if (in.filter == FilterNearest) {
        int ix, iy, iz;
        nearest_index(in, p, ix, iy, iz);
        return vol_at(in, vb, ix, iy, iz);
    }
    TriCell<T> c;
    trilinear_cell(in, vb, p, c);
    return trilinear_value(c);
}

// Pre-shaded volume (renderer_interpolation.cuh:390-430): the four batch slices
// of the volume hold r, g, b and absorption; each is interpolated trilinearly
// and the TF passes the colour on (renderer_tf.cuh:818-854).
template <class T>
inline V4<T> preshaded_fetch(const Inputs<T>& in, V3<T> p, TriCell<T> cells[4])
{
    T v[4];
    for (int ch = 0; ch < 4; ++ch) {
        trilinear_cell(in, ch, p, cells[ch]);
        v[ch] = trilinear_value(cells[ch]);
    }
    return v4(v[0], v[1], v[2], v[3]);
}

// Trilinear adjoint (:360-387): per-corner weight products; adj_p = adj_cw - adj_fw.
// Corner adjoints are returned in w[8] (same bit order as TriCell::d).
template <class T>
inline V3<T> trilinear_adjoint(const TriCell<T>& c, T adj_d, T w[8])
{
    const V3<T> cw = v3(c.fx, c.fy, c.fz);
    const V3<T> fw = v3(T(1) - c.fx, T(1) - c.fy, T(1) - c.fz);
    V3<T> acw = v3(T(0), T(0), T(0)), afw = v3(T(0), T(0), T(0));
    for (int k = 0; k < 8; ++k) {
        const bool bx = (k >> 2) & 1, by = (k >> 1) & 1, bz = k & 1;
        const T X = bx ? cw.x : fw.x, Y = by ? cw.y : fw.y, Z = bz ? cw.z : fw.z;
        w[k] = adj_d * (X * Y * Z);
        const T gx = adj_d * c.d[k] * Y * Z;
        const T gy = adj_d * c.d[k] * X * Z;
        const T gz = adj_d * c.d[k] * X * Y;
        if (bx) acw.x += gx; else afw.x += gx;
        if (by) acw.y += gy; else afw.y += gy;
        if (bz) acw.z += gz; else afw.z += gz;
    }
    return acw - afw;
}

// ---------------------------------------------------------------------------
// a5/a9: transfer functions (renderer_tf.cuh:59-648)
// ---------------------------------------------------------------------------
template <class T> inline const T* tf_row(const Inputs<T>& in, int b, int r)
{
    return in.tf + ((size_t)bidx(in.tf_b, b) * in.R + r) * in.C;
}

template <class T>
inline V4<T> tf_eval(const Inputs<T>& in, int b, T density)
{
    switch (in.tf_mode) {
    case TFIdentity: {
        const T* t = tf_row(in, b, 0);
        const T d = fclamp(density, T(0), T(1));
        return v4(d * t[1], d * t[1], d * t[1], d * t[0]);
    }
    case TFTexture: {
        const int R = in.R;
        const T d = density * T(R) - T(0.5);
        const int di = (int)std::floor(d);
        const T df = d - T(di);
        const T* a = tf_row(in, b, iclamp(di, 0, R - 1));
        const T* c = tf_row(in, b, iclamp(di + 1, 0, R - 1));
        return v4(lerp_(a[0], c[0], df), lerp_(a[1], c[1], df), lerp_(a[2], c[2], df), lerp_(a[3], c[3], df));
    }
    case TFLinear: {
        const int R = in.R;
        int i;
        for (i = 0; i < R - 2; ++i)
            if (tf_row(in, b, i + 1)[4] > density) break;
        const T* a = tf_row(in, b, i);
        const T* c = tf_row(in, b, i + 1);
        const T p0 = a[4], p1 = c[4];
        const T d2 = fclamp(density, p0, p1);
        const T f = (d2 - p0) / (p1 - p0);
        return v4(lerp_(a[0], c[0], f), lerp_(a[1], c[1], f), lerp_(a[2], c[2], f), lerp_(a[3], c[3], f));
    }
    case TFGaussianLog: {  // renderer_tf.cuh:676-688: log-sum-exp over i of log c_i - (d-mu_i)^2/(2 sigma_i^2)
        const int R = in.R;
        T m[4], sum[4] = {T(0), T(0), T(0), T(0)};
        for (int ch = 0; ch < 4; ++ch) m[ch] = -std::numeric_limits<T>::infinity();
        for (int pass = 0; pass < 2; ++pass)
            for (int i = 0; i < R; ++i) {
                const T* r = tf_row(in, b, i);
                const T ln = -(density - r[4]) * (density - r[4]) / (T(2) * r[5] * r[5]);
                for (int ch = 0; ch < 4; ++ch) {
                    const T t = std::log(r[ch]) + ln;
                    if (pass == 0) m[ch] = std::max(m[ch], t);
                    else sum[ch] += std::exp(t - m[ch]);
                }
            }
        return v4(m[0] + std::log(sum[0]), m[1] + std::log(sum[1]), m[2] + std::log(sum[2]), m[3] + std::log(sum[3]));
    }
    default: {  // TFGaussian: sum_i c_i exp(-(d-mu)^2/(2 sigma^2)), column 5 is sigma
        V4<T> out = v4(T(0), T(0), T(0), T(0));
        for (int i = 0; i < in.R; ++i) {
            const T* r = tf_row(in, b, i);
            const T mu = r[4], s = r[5];
            const T n = std::exp(-(density - mu) * (density - mu) / (T(2) * s * s));
            out.x += r[0] * n; out.y += r[1] * n; out.z += r[2] * n; out.w += r[3] * n;
        }
        return out;
    }
    }
}

// TF adjoint.  adj_tf (may be null) receives the parameter gradient for batch tb.
template <class T>
inline T tf_adjoint(const Inputs<T>& in, int b, T density, V4<T> adj, T* adj_tf /*R x C of this batch or null*/)
{
    const int C = in.C;
    switch (in.tf_mode) {
    case TFIdentity: {  // :75-104
        const T* t = tf_row(in, b, 0);
        const T dc = fclamp(density, T(0), T(1));
        const T adj_dc = (adj.x * t[1] + adj.y * t[1] + adj.z * t[1]) + adj.w * t[0];
        if (adj_tf) {
            atomic_add(&adj_tf[0], adj.w * dc);
            atomic_add(&adj_tf[1], adj.x * dc + adj.y * dc + adj.z * dc);
        }
        return (density > T(0) && density < T(1)) ? adj_dc : T(0);
    }
    case TFTexture: {  // :173-220
        const int R = in.R;
        const T d = density * T(R) - T(0.5);
        const int di = (int)std::floor(d);
        const T df = d - T(di);
        const int i0 = iclamp(di, 0, R - 1), i1 = iclamp(di + 1, 0, R - 1);
        const T* a = tf_row(in, b, i0);
        const T* c = tf_row(in, b, i1);
        const T adj_df = adj.x * (c[0] - a[0]) + adj.y * (c[1] - a[1]) + adj.z * (c[2] - a[2]) + adj.w * (c[3] - a[3]);
        if (adj_tf) {
            const T w0 = T(1) - df;
            atomic_add(&adj_tf[i0 * C + 0], adj.x * w0); atomic_add(&adj_tf[i0 * C + 1], adj.y * w0);
            atomic_add(&adj_tf[i0 * C + 2], adj.z * w0); atomic_add(&adj_tf[i0 * C + 3], adj.w * w0);
            atomic_add(&adj_tf[i1 * C + 0], adj.x * df); atomic_add(&adj_tf[i1 * C + 1], adj.y * df);
            atomic_add(&adj_tf[i1 * C + 2], adj.z * df); atomic_add(&adj_tf[i1 * C + 3], adj.w * df);
        }
        return adj_df * T(R);
    }
    case TFLinear: {  // :315-400 (gradient passes through the clamp, :350-351)
        const int R = in.R;
        int i;
        for (i = 0; i < R - 2; ++i)
            if (tf_row(in, b, i + 1)[4] > density) break;
        const T* a = tf_row(in, b, i);
        const T* c = tf_row(in, b, i + 1);
        const T p0 = a[4], p1 = c[4];
        const T d2 = fclamp(density, p0, p1);
        const T f = (d2 - p0) / (p1 - p0);
        const T adj_f = adj.x * (c[0] - a[0]) + adj.y * (c[1] - a[1]) + adj.z * (c[2] - a[2]) + adj.w * (c[3] - a[3]);
        const T den = (p1 - p0) * (p1 - p0);
        const T adj_p0 = adj_f * (d2 - p1) / den;
        const T adj_p1 = adj_f * (p0 - d2) / den;
        const T adj_d2 = adj_f / (p1 - p0);
        if (adj_tf) {
            const T w0 = T(1) - f;
            atomic_add(&adj_tf[i * C + 0], adj.x * w0); atomic_add(&adj_tf[i * C + 1], adj.y * w0);
            atomic_add(&adj_tf[i * C + 2], adj.z * w0); atomic_add(&adj_tf[i * C + 3], adj.w * w0);
            atomic_add(&adj_tf[i * C + 4], adj_p0);
            const int j = i + 1;
            atomic_add(&adj_tf[j * C + 0], adj.x * f); atomic_add(&adj_tf[j * C + 1], adj.y * f);
            atomic_add(&adj_tf[j * C + 2], adj.z * f); atomic_add(&adj_tf[j * C + 3], adj.w * f);
            atomic_add(&adj_tf[j * C + 4], adj_p1);
        }
        return adj_d2;
    }
    case TFGaussianLog: {  // :706-744, the density term (adj_tf is not formed: pre-shading does not ask for it)
        const V4<T> lse = tf_eval(in, b, density);
        const T l[4] = {lse.x, lse.y, lse.z, lse.w}, a[4] = {adj.x, adj.y, adj.z, adj.w};
        T adj_d = T(0);
        for (int i = 0; i < in.R; ++i) {
            const T* r = tf_row(in, b, i);
            const T mu = r[4], s2 = r[5] * r[5];
            const T ln = -(density - mu) * (density - mu) / (T(2) * s2);
            for (int ch = 0; ch < 4; ++ch) adj_d += a[ch] * std::exp(std::log(r[ch]) + ln - l[ch]) * (mu - density) / s2;
        }
        return adj_d;
    }
    default: {  // TFGaussian :562-614
        T adj_d = T(0);
        for (int i = 0; i < in.R; ++i) {
            const T* r = tf_row(in, b, i);
            const V4<T> ci = v4(r[0], r[1], r[2], r[3]);
            const T mu = r[4], s = r[5];
            const T s2 = s * s;
            const T n = std::exp(-(density - mu) * (density - mu) / (T(2) * s * s));
            const T f1 = (mu - density) / s2 * n;
            adj_d += dot(adj, v4(ci.x * f1, ci.y * f1, ci.z * f1, ci.w * f1));
            if (adj_tf) {
                const T f2 = f1 * (mu - density) / s;
                atomic_add(&adj_tf[i * C + 0], adj.x * n); atomic_add(&adj_tf[i * C + 1], adj.y * n);
                atomic_add(&adj_tf[i * C + 2], adj.z * n); atomic_add(&adj_tf[i * C + 3], adj.w * n);
                atomic_add(&adj_tf[i * C + 4], -dot(adj, v4(ci.x * f1, ci.y * f1, ci.z * f1, ci.w * f1)));
                atomic_add(&adj_tf[i * C + 5], dot(adj, v4(ci.x * f2, ci.y * f2, ci.z * f2, ci.w * f2)));
            }
        }
        return adj_d;
    }
    }
}

// ---------------------------------------------------------------------------
// a6/a8: blending (renderer_blending.cuh:36-165)
// ---------------------------------------------------------------------------
template <class T> inline T blend_alpha(int mode, T absorption, T step)
{
    if (mode == BlendBeerLambert) return T(1) - std::exp(-absorption * step);
    return tmin_(T(1), absorption * step);
}

template <class T> inline V4<T> blend(int mode, V4<T> acc, V4<T> cur, T step)
{
    const T a = blend_alpha(mode, cur.w, step);
    const T tr = T(1) - acc.w;
    return v4(acc.x + tr * cur.x * a, acc.y + tr * cur.y * a, acc.z + tr * cur.z * a, acc.w + tr * a);
}

// Blending adjoint given the accumulator alpha *before* the sample (a_in).
// Returns adj of the sample (rgb, absorption) and of the step; updates
// adj_acc in place (rgb part unchanged, alpha part propagated).
template <class T>
inline void blend_adjoint_given(int mode, T a_in, V4<T> cur, T step, V4<T>& adj_acc, V4<T>& adj_cur, T& adj_step)
{
    T alpha;
    if (mode == BlendBeerLambert) alpha = T(1) - std::exp(-cur.w * step);
    else alpha = tmin_(T(1) - T(1e-5), cur.w * step);
    T adj_alpha = adj_acc.w * (T(1) - a_in);
    T adj_ain = adj_acc.w * (T(1) - alpha);
    adj_alpha += adj_acc.x * (cur.x - cur.x * a_in) + adj_acc.y * (cur.y - cur.y * a_in) + adj_acc.z * (cur.z - cur.z * a_in);
    adj_ain += adj_acc.x * (-cur.x * alpha) + adj_acc.y * (-cur.y * alpha) + adj_acc.z * (-cur.z * alpha);
    const T s = alpha * (T(1) - a_in);
    adj_cur = v4(adj_acc.x * s, adj_acc.y * s, adj_acc.z * s, T(0));
    if (mode == BlendBeerLambert) {
        const T e = std::exp(-cur.w * step);
        adj_cur.w = adj_alpha * step * e;
        adj_step = adj_alpha * cur.w * e;
    } else {
        const bool lin = cur.w * step < T(1);
        adj_step = lin ? adj_alpha * cur.w : T(0);
        adj_cur.w = lin ? adj_alpha * step : T(0);
    }
    adj_acc.w = adj_ain;
}

// Inversion (reference semantics, :63-98 / :128-164): reconstruct the
// accumulator before the sample from the one after it.
template <class T> inline V4<T> blend_invert(int mode, V4<T> acc_out, V4<T> cur, T step)
{
    T alpha;
    if (mode == BlendBeerLambert) alpha = T(1) - std::exp(-cur.w * step);
    else alpha = tmin_(T(1) - T(1e-5), cur.w * step);
    const T a_in = (alpha - acc_out.w) / (alpha - T(1));
    const T tr = T(1) - a_in;
    return v4(acc_out.x - tr * cur.x * alpha, acc_out.y - tr * cur.y * alpha, acc_out.z - tr * cur.z * alpha, a_in);
}

// ---------------------------------------------------------------------------
// Per-ray setup shared by forward and adjoint (renderer_kernels.cuh:31-48)
// ---------------------------------------------------------------------------
template <class T> struct Ray {
    V3<T> start, dir, bmin, bsize, voxel;
    T step, t0, t1;
};

template <class T>
inline Ray<T> setup_ray(const Inputs<T>& in, int x, int y, int b)
{
    Ray<T> r;
    const T* bm = in.box_min + 3 * bidx(in.box_min_b, b);
    const T* bs = in.box_size + 3 * bidx(in.box_size_b, b);
    r.bmin = v3(bm[0], bm[1], bm[2]);
    r.bsize = v3(bs[0], bs[1], bs[2]);
    r.voxel = r.bsize / v3(T(in.X - 1), T(in.Y - 1), T(in.Z - 1));
    camera_rays(in, x, y, b, r.start, r.dir);
    const int sb = bidx(in.step_b, b), sy = in.step_h == 1 ? 0 : y, sx = in.step_w == 1 ? 0 : x;
    r.step = in.step[((size_t)sb * in.step_h + sy) * in.step_w + sx];
    ray_aabb(r.start, r.dir, r.bmin, r.bsize, r.t0, r.t1);
    r.t0 = tmax_(r.t0, T(0));
    // degenerate rays / steps sample nothing (the reference would not terminate)
    if (!(r.t1 >= r.t0) || !std::isfinite(r.t0) || !std::isfinite(r.t1) || !(r.step > T(0)) || !std::isfinite(r.step)) {
        r.t0 = T(0);
        r.t1 = T(-1);
        if (!(r.step > T(0)) || !std::isfinite(r.step)) r.step = T(1);
    }
    return r;
}

template <class T> inline V3<T> sample_pos(const Ray<T>& r, T t)
{
    const V3<T> w = r.start + t * r.dir;
    return (w - r.bmin) / r.voxel;
}

// ---------------------------------------------------------------------------
// a1: forward march for one ray.  frozen_n >= 0 fixes the sample count
// (ignoring t>tmax and the early-out), frozen_t0 (if non-null) fixes tmin;
// both are used by the finite-difference protocol.
// ---------------------------------------------------------------------------
template <class T>
inline void march_forward(const Inputs<T>& in, int x, int y, int b, int frozen_n, const T* frozen_t0,
                          V4<T>& color, int& term, T* t0_out, std::vector<T>* a_in_trace)
{
    Ray<T> r = setup_ray(in, x, y, b);
    if (frozen_t0) r.t0 = *frozen_t0;
    if (t0_out) *t0_out = r.t0;
    V4<T> acc = v4(T(0), T(0), T(0), T(0));
    const int vb = bidx(in.vol_b, b);
    int i;
    // cell cache as in renderer_interpolation.cuh:218-225 (does not change values)
    TriCell<T> cell{};
    cell.ix = cell.iy = cell.iz = -1 - (1 << 30);
    for (i = 0; i < (1 << 24); ++i) {
        const T t = r.t0 + T(i) * r.step;
        if (frozen_n >= 0) {
            if (i >= frozen_n) break;
        } else {
            if (t > r.t1) break;
            if (acc.w > in.early_out) break;
        }
        const V3<T> p = sample_pos(r, t);
        V4<T> s;
        if (in.filter == FilterPreshaded) {
            TriCell<T> cells[4];
            s = preshaded_fetch(in, p, cells);
        } else {
            T d;
            if (in.filter == FilterNearest) {
                d = fetch_density(in, b, p);
            } else {
                const int ix = (int)p.x, iy = (int)p.y, iz = (int)p.z;
                if (ix != cell.ix || iy != cell.iy || iz != cell.iz) trilinear_cell(in, vb, p, cell);
                cell.fx = p.x - T(ix); cell.fy = p.y - T(iy); cell.fz = p.z - T(iz);
                d = trilinear_value(cell);
            }
            s = tf_eval(in, b, d);
        }
        if (a_in_trace) a_in_trace->push_back(acc.w);
        acc = blend(in.blend, acc, s, r.step);
    }
    color = acc;
    term = i;
}

// ---------------------------------------------------------------------------
// a7/a11: adjoint march for one ray (renderer_kernels.cuh:324-444).
//   method == AdjointInversion : reference semantics (reconstruct by inversion)
//   method == AdjointExactReplay: replay the forward with the same sample count
//       and keep every pre-sample accumulator (no inversion).
// ---------------------------------------------------------------------------
template <class T>
inline void march_adjoint(const Inputs<T>& in, int x, int y, int b, V4<T> color_out, int term, V4<T> adj_color,
                          const AdjointOutputs<T>& out, int method, const T* frozen_t0)
{
    Ray<T> r = setup_ray(in, x, y, b);
    if (frozen_t0) r.t0 = *frozen_t0;
    std::vector<T> a_in;
    if (method == AdjointExactReplay) {
        V4<T> c; int n;
        a_in.reserve(term > 0 ? term : 1);
        march_forward(in, x, y, b, term, frozen_t0, c, n, (T*)nullptr, &a_in);
    }
    const int vb = bidx(in.vol_b, b);
    T* adj_tf = out.adj_tf ? out.adj_tf + (size_t)bidx(out.adj_tf_b, b) * in.R * in.C : nullptr;
    T* adj_vol = out.adj_volume ? out.adj_volume + (size_t)bidx(out.adj_volume_b, b) * in.X * in.Y * in.Z : nullptr;
    V4<T> acc = color_out;
    V4<T> adj_acc = adj_color;
    T adj_step = T(0);
    V3<T> adj_start = v3(T(0), T(0), T(0)), adj_dir = v3(T(0), T(0), T(0));
    for (int i = term - 1; i >= 0; --i) {
        const T t = r.t0 + T(i) * r.step;
        const V3<T> p = sample_pos(r, t);
        TriCell<T> cell{};
        TriCell<T> cells[4];
        T d = T(0);
        int nx = 0, ny = 0, nz = 0;
        V4<T> s;
        if (in.filter == FilterPreshaded) {
            s = preshaded_fetch(in, p, cells);
        } else {
            if (in.filter == FilterNearest) {
                nearest_index(in, p, nx, ny, nz);
                d = vol_at(in, vb, nx, ny, nz);
            } else {
                trilinear_cell(in, vb, p, cell);
                d = trilinear_value(cell);
            }
            s = tf_eval(in, b, d);
        }
        T ain;
        if (method == AdjointExactReplay) {
            ain = a_in[(size_t)i];
        } else {
            acc = blend_invert(in.blend, acc, s, r.step);
            ain = acc.w;
        }
        V4<T> adj_s; T adj_step_local;
        blend_adjoint_given(in.blend, ain, s, r.step, adj_acc, adj_s, adj_step_local);
        V3<T> adj_p = v3(T(0), T(0), T(0));
        if (in.filter == FilterPreshaded) {
            // the gradient of each channel's trilinear form (the reference puts the z term of this
            // sum into y, renderer_interpolation.cuh:557-562)
            const T as[4] = {adj_s.x, adj_s.y, adj_s.z, adj_s.w};
            for (int ch = 0; ch < 4; ++ch) {
                T w[8];
                adj_p = adj_p + trilinear_adjoint(cells[ch], as[ch], w);
                if (adj_vol) {
                    T* av = out.adj_volume + (size_t)ch * in.X * in.Y * in.Z;
                    for (int k = 0; k < 8; ++k) {
                        const int ax = (k >> 2) & 1 ? cells[ch].hx : cells[ch].lx;
                        const int ay = (k >> 1) & 1 ? cells[ch].hy : cells[ch].ly;
                        const int az = k & 1 ? cells[ch].hz : cells[ch].lz;
                        atomic_add(&av[((size_t)ax * in.Y + ay) * in.Z + az], w[k]);
                    }
                }
            }
        }
        const T adj_d = in.filter == FilterPreshaded ? T(0) : tf_adjoint(in, b, d, adj_s, adj_tf);
        if (in.filter == FilterPreshaded) {
        } else if (in.filter == FilterNearest) {
            if (adj_vol) atomic_add(&adj_vol[((size_t)nx * in.Y + ny) * in.Z + nz], adj_d);
        } else {
            T w[8];
            adj_p = trilinear_adjoint(cell, adj_d, w);
            if (adj_vol) {
                for (int k = 0; k < 8; ++k) {
                    const int ax = (k >> 2) & 1 ? cell.hx : cell.lx;
                    const int ay = (k >> 1) & 1 ? cell.hy : cell.ly;
                    const int az = k & 1 ? cell.hz : cell.lz;
                    atomic_add(&adj_vol[((size_t)ax * in.Y + ay) * in.Z + az], w[k]);
                }
            }
        }
        const V3<T> adj_w = adj_p / r.voxel;
        adj_start = adj_start + adj_w;
        adj_dir = adj_dir + adj_w * t;
        adj_step += adj_step_local + dot(adj_w, r.dir) * T(i);
    }
    if (out.adj_step) {
        const int sb = bidx(out.adj_step_b, b), sy = out.adj_step_h == 1 ? 0 : y, sx = out.adj_step_w == 1 ? 0 : x;
        atomic_add(&out.adj_step[((size_t)sb * out.adj_step_h + sy) * out.adj_step_w + sx], adj_step);
    }
    if (out.adj_ray_start) {
        const size_t o = (((size_t)bidx(out.adj_cam_b, b) * in.H + y) * in.W + x) * 3;
        atomic_add(&out.adj_ray_start[o + 0], adj_start.x);
        atomic_add(&out.adj_ray_start[o + 1], adj_start.y);
        atomic_add(&out.adj_ray_start[o + 2], adj_start.z);
        atomic_add(&out.adj_ray_dir[o + 0], adj_dir.x);
        atomic_add(&out.adj_ray_dir[o + 1], adj_dir.y);
        atomic_add(&out.adj_ray_dir[o + 2], adj_dir.z);
    }
}

// ---------------------------------------------------------------------------
// a12: camera utilities (renderer_adjoint.cuh:28-65; camera_impl.cu:17-174)
// ---------------------------------------------------------------------------
template <class T> inline V3<T> adj_normalize(V3<T> in, V3<T> adj_out)
{
    const T len = length(in);
    V3<T> adj_in = adj_out / len;
    const T adj_len = dot(adj_out, (T(-1) * in) / (len * len));
    return adj_in + (adj_len * in) / len;
}

template <class T> inline void adj_cross(V3<T> a, V3<T> b, V3<T> adj_c, V3<T>& adj_a, V3<T>& adj_b)
{
    adj_a.y += b.z * adj_c.x; adj_b.z += a.y * adj_c.x; adj_a.z -= b.y * adj_c.x; adj_b.y -= a.z * adj_c.x;
    adj_a.z += b.x * adj_c.y; adj_b.x += a.z * adj_c.y; adj_a.x -= b.z * adj_c.y; adj_b.z -= a.x * adj_c.y;
    adj_a.x += b.y * adj_c.z; adj_b.y += a.x * adj_c.z; adj_a.y -= b.x * adj_c.z; adj_b.x -= a.y * adj_c.z;
}

// generateRays forward: viewport B x 3 x 3 (eye,right,up) -> B x H x W x 3 twice.
// The reference evaluates tan in float (camera_impl.cu:74-76); the oracle uses T.
template <class T>
inline void generate_rays(const T* viewport, int B, int H, int W, T fov_y, T* ray_start, T* ray_dir)
{
    const T fovX = fov_y * T(W) / T(H);
    const T tanX = std::tan(fovX / T(2)), tanY = std::tan(fov_y / T(2));
#pragma omp parallel for schedule(static)
    for (long long idx = 0; idx < (long long)B * H * W; ++idx) {
        const int b = (int)(idx / ((long long)H * W));
        const int y = (int)((idx / W) % H);
        const int x = (int)(idx % W);
        const T* m = viewport + 9 * b;
        const V3<T> eye = v3(m[0], m[1], m[2]), right = v3(m[3], m[4], m[5]), up = v3(m[6], m[7], m[8]);
        const V3<T> front = cross(up, right);
        const T fx = T(2) * (T(x) + T(0.5)) / T(W) - T(1);
        const T fy = T(2) * (T(y) + T(0.5)) / T(H) - T(1);
        const V3<T> d = normalize(front + (fx * tanX) * right + (fy * tanY) * up);
        ray_start[3 * idx + 0] = eye.x; ray_start[3 * idx + 1] = eye.y; ray_start[3 * idx + 2] = eye.z;
        ray_dir[3 * idx + 0] = d.x; ray_dir[3 * idx + 1] = d.y; ray_dir[3 * idx + 2] = d.z;
    }
}

// generateRays adjoint: per-pixel adjNormalize/adjCross, summed per camera
// (camera_impl.cu:134-174, 302-357).  adj_viewport is overwritten.
template <class T>
inline void generate_rays_adjoint(const T* viewport, int B, int H, int W, T fov_y, const T* adj_start,
                                  const T* adj_dir, T* adj_viewport)
{
    const T fovX = fov_y * T(W) / T(H);
    const T tanX = std::tan(fovX / T(2)), tanY = std::tan(fov_y / T(2));
#pragma omp parallel for schedule(static)
    for (int b = 0; b < B; ++b) {
        const T* m = viewport + 9 * b;
        const V3<T> right = v3(m[3], m[4], m[5]), up = v3(m[6], m[7], m[8]);
        const V3<T> front = cross(up, right);
        V3<T> se = v3(T(0), T(0), T(0)), sr = se, su = se;
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                const size_t o = (((size_t)b * H + y) * W + x) * 3;
                const T fx = T(2) * (T(x) + T(0.5)) / T(W) - T(1);
                const T fy = T(2) * (T(y) + T(0.5)) / T(H) - T(1);
                const V3<T> d = front + (fx * tanX) * right + (fy * tanY) * up;
                const V3<T> ae = v3(adj_start[o], adj_start[o + 1], adj_start[o + 2]);
                const V3<T> adn = v3(adj_dir[o], adj_dir[o + 1], adj_dir[o + 2]);
                const V3<T> ad = adj_normalize(d, adn);
                V3<T> ar = ad * (fx * tanX);
                V3<T> au = ad * (fy * tanY);
                adj_cross(up, right, ad, au, ar);
                se = se + ae; sr = sr + ar; su = su + au;
            }
        T* o = adj_viewport + 9 * b;
        o[0] = se.x; o[1] = se.y; o[2] = se.z; o[3] = sr.x; o[4] = sr.y; o[5] = sr.z; o[6] = su.x; o[7] = su.y; o[8] = su.z;
    }
}

// ---------------------------------------------------------------------------
// Whole-image drivers (renderer_kernels.cuh:114-127, 480-496): one ray per
// loop iteration over B*H*W, x fastest, OpenMP guided schedule.
// ---------------------------------------------------------------------------
template <class T>
inline void render_forward(const Inputs<T>& in, T* color, int* term, const int* frozen_term, const T* frozen_t0, T* t0_out)
{
    const long long count = (long long)in.B * in.H * in.W;
#pragma omp parallel for schedule(guided)
    for (long long i = 0; i < count; ++i) {
        const int b = (int)(i / ((long long)in.W * in.H));
        const int y = (int)((i - (long long)b * in.W * in.H) / in.W);
        const int x = (int)(i - (long long)in.W * (y + (long long)in.H * b));
        V4<T> c; int n;
        march_forward(in, x, y, b, frozen_term ? frozen_term[i] : -1, frozen_t0 ? frozen_t0 + i : (const T*)nullptr,
                      c, n, t0_out ? t0_out + i : (T*)nullptr, (std::vector<T>*)nullptr);
        color[4 * i + 0] = c.x; color[4 * i + 1] = c.y; color[4 * i + 2] = c.z; color[4 * i + 3] = c.w;
        term[i] = n;
    }
}

template <class T>
inline void render_adjoint(const Inputs<T>& in, const T* color, const int* term, const T* adj_color,
                           const AdjointOutputs<T>& out, int method, const T* frozen_t0)
{
    const long long count = (long long)in.B * in.H * in.W;
#pragma omp parallel for schedule(guided)
    for (long long i = 0; i < count; ++i) {
        const int b = (int)(i / ((long long)in.W * in.H));
        const int y = (int)((i - (long long)b * in.W * in.H) / in.W);
        const int x = (int)(i - (long long)in.W * (y + (long long)in.H * b));
        march_adjoint(in, x, y, b, v4(color[4 * i], color[4 * i + 1], color[4 * i + 2], color[4 * i + 3]), term[i],
                      v4(adj_color[4 * i], adj_color[4 * i + 1], adj_color[4 * i + 2], adj_color[4 * i + 3]), out,
                      method, frozen_t0 ? frozen_t0 + i : (const T*)nullptr);
    }
}

// ---------------------------------------------------------------------------
// Forward-mode differentiation (renderer_kernels.cuh:135-308,
// renderer_tf.cuh:122-145,256-287,437-485,522-560): every traced quantity is
// a dual number carrying D derivatives, as the reference's cudAD
// fvar<real_t, D>.  Traced inputs: step size, the evaluated ray start and
// direction, TF entries (index table d_tf, -1 = not traced).  tmin is not
// traced (the reference casts the ray to real for the slab test).
//
// cudAD is not available, so its clamp convention is taken from the
// reference's adjoint, which the reference's consistency tests hold the
// forward mode to (unittests/testConsistencyTF.cpp:183-204):
//   identity TF : density derivative passes strictly inside (0,1) only
//                 (renderer_tf.cuh:103);
//   linear TF   : clamp(d, p0, p1) keeps the derivative of d, the fraction is
//                 differentiated with the clamped value (renderer_tf.cuh:337-357);
//   alpha blend : min(1, x) passes the derivative of x iff x < 1
//                 (renderer_blending.cuh:118-125 with the adjoint :147-159).
// ---------------------------------------------------------------------------
template <class T> struct Dual {
    T v;
    std::vector<T> d;
};
template <class T> inline Dual<T> dconst(T v, int D) { return Dual<T>{v, std::vector<T>((size_t)D, T(0))}; }
template <class T> inline Dual<T> dinput(T v, int idx, int D)
{
    Dual<T> r = dconst(v, D);
    if (idx >= 0) r.d[(size_t)idx] = T(1);
    return r;
}
template <class T> inline Dual<T> operator+(const Dual<T>& a, const Dual<T>& b)
{
    Dual<T> r{a.v + b.v, a.d};
    for (size_t k = 0; k < r.d.size(); ++k) r.d[k] += b.d[k];
    return r;
}
template <class T> inline Dual<T> operator-(const Dual<T>& a, const Dual<T>& b)
{
    Dual<T> r{a.v - b.v, a.d};
    for (size_t k = 0; k < r.d.size(); ++k) r.d[k] -= b.d[k];
    return r;
}
template <class T> inline Dual<T> operator*(const Dual<T>& a, const Dual<T>& b)
{
    Dual<T> r{a.v * b.v, a.d};
    for (size_t k = 0; k < r.d.size(); ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
template <class T> inline Dual<T> operator/(const Dual<T>& a, const Dual<T>& b)
{
    Dual<T> r{a.v / b.v, a.d};
    for (size_t k = 0; k < r.d.size(); ++k) r.d[k] = (a.d[k] * b.v - a.v * b.d[k]) / (b.v * b.v);
    return r;
}
template <class T> inline Dual<T> operator*(T s, const Dual<T>& a)
{
    Dual<T> r{s * a.v, a.d};
    for (auto& x : r.d) x *= s;
    return r;
}
template <class T> inline Dual<T> operator-(const Dual<T>& a, T s) { return Dual<T>{a.v - s, a.d}; }
template <class T> inline Dual<T> operator+(const Dual<T>& a, T s) { return Dual<T>{a.v + s, a.d}; }
template <class T> inline Dual<T> operator-(T s, const Dual<T>& a) { return T(-1) * a + s; }
template <class T> inline Dual<T> operator/(const Dual<T>& a, T s)
{
    Dual<T> r{a.v / s, a.d};
    for (auto& x : r.d) x /= s;
    return r;
}
template <class T> inline Dual<T> dexp(const Dual<T>& a)
{
    const T e = std::exp(a.v);
    Dual<T> r{e, a.d};
    for (auto& x : r.d) x *= e;
    return r;
}
template <class T> inline Dual<T> dlerp(const Dual<T>& a, const Dual<T>& b, const Dual<T>& t) { return a + t * (b - a); }

struct FwdDiffSettings {
    int D;
    int d_step;
    int d_start[3], d_dir[3];
    const int* d_tf;  // d_tf_b x R x C, or null
    int d_tf_b;
};

template <class T>
inline Dual<T> tf_entry(const Inputs<T>& in, const FwdDiffSettings& fd, int b, int r, int c)
{
    const int idx = fd.d_tf ? fd.d_tf[((size_t)bidx(fd.d_tf_b, b) * in.R + r) * in.C + c] : -1;
    return dinput(tf_row(in, b, r)[c], idx, fd.D);
}

template <class T>
inline void tf_eval_dual(const Inputs<T>& in, const FwdDiffSettings& fd, int b, const Dual<T>& density,
                         Dual<T> out[4])
{
    const int D = fd.D;
    switch (in.tf_mode) {
    case TFIdentity: {
        Dual<T> dc{fclamp(density.v, T(0), T(1)), density.d};
        if (!(density.v > T(0) && density.v < T(1))) dc.d.assign((size_t)D, T(0));
        const Dual<T> sa = tf_entry(in, fd, b, 0, 0), sc = tf_entry(in, fd, b, 0, 1);
        out[0] = dc * sc; out[1] = dc * sc; out[2] = dc * sc; out[3] = dc * sa;
        return;
    }
    case TFTexture: {
        const int R = in.R;
        const Dual<T> d = T(R) * density - T(0.5);
        const int di = (int)std::floor(d.v);
        const Dual<T> df = d - T(di);
        const int r0 = iclamp(di, 0, R - 1), r1 = iclamp(di + 1, 0, R - 1);
        for (int c = 0; c < 4; ++c) out[c] = dlerp(tf_entry(in, fd, b, r0, c), tf_entry(in, fd, b, r1, c), df);
        return;
    }
    case TFLinear: {
        const int R = in.R;
        int i;
        for (i = 0; i < R - 2; ++i)
            if (tf_row(in, b, i + 1)[4] > density.v) break;
        const Dual<T> p0 = tf_entry(in, fd, b, i, 4), p1 = tf_entry(in, fd, b, i + 1, 4);
        const Dual<T> d2{fclamp(density.v, p0.v, p1.v), density.d};
        const Dual<T> f = (d2 - p0) / (p1 - p0);
        for (int c = 0; c < 4; ++c) out[c] = dlerp(tf_entry(in, fd, b, i, c), tf_entry(in, fd, b, i + 1, c), f);
        return;
    }
    default: {
        for (int c = 0; c < 4; ++c) out[c] = dconst(T(0), D);
        for (int i = 0; i < in.R; ++i) {
            const Dual<T> mu = tf_entry(in, fd, b, i, 4), s = tf_entry(in, fd, b, i, 5);
            const Dual<T> dm = density - mu;
            const Dual<T> n = dexp(T(-1) * (dm * dm) / (T(2) * (s * s)));
            for (int c = 0; c < 4; ++c) out[c] = out[c] + tf_entry(in, fd, b, i, c) * n;
        }
        return;
    }
    }
}

// One ray; grad receives D x 4 derivatives of the RGBA output.
template <class T>
inline void march_forward_gradients(const Inputs<T>& in, const FwdDiffSettings& fd, int x, int y, int b,
                                    int frozen_n, const T* frozen_t0, V4<T>& color, int& term, T* grad)
{
    const int D = fd.D;
    Ray<T> r = setup_ray(in, x, y, b);
    if (frozen_t0) r.t0 = *frozen_t0;
    const T sv[3] = {r.start.x, r.start.y, r.start.z}, dv[3] = {r.dir.x, r.dir.y, r.dir.z};
    const T bm[3] = {r.bmin.x, r.bmin.y, r.bmin.z}, vx[3] = {r.voxel.x, r.voxel.y, r.voxel.z};
    std::vector<Dual<T>> start, dir;
    for (int c = 0; c < 3; ++c) {
        start.push_back(dinput(sv[c], fd.d_start[c], D));
        dir.push_back(dinput(dv[c], fd.d_dir[c], D));
    }
    const Dual<T> step = dinput(r.step, fd.d_step, D);
    Dual<T> acc[4] = {dconst(T(0), D), dconst(T(0), D), dconst(T(0), D), dconst(T(0), D)};
    const int vb = bidx(in.vol_b, b);
    int i;
    for (i = 0; i < (1 << 24); ++i) {
        const Dual<T> t = T(i) * step + r.t0;
        if (frozen_n >= 0) {
            if (i >= frozen_n) break;
        } else {
            if (t.v > r.t1) break;
            if (acc[3].v > in.early_out) break;
        }
        Dual<T> p[3];
        for (int c = 0; c < 3; ++c) p[c] = ((start[c] + t * dir[c]) - bm[c]) / vx[c];
        Dual<T> density;
        if (in.filter == FilterNearest) {
            density = dconst(fetch_density(in, b, v3(p[0].v, p[1].v, p[2].v)), D);
        } else {
            TriCell<T> cell;
            trilinear_cell(in, vb, v3(p[0].v, p[1].v, p[2].v), cell);
            const Dual<T> fx = p[0] - T(cell.ix), fy = p[1] - T(cell.iy), fz = p[2] - T(cell.iz);
            Dual<T> cd[8];
            for (int k = 0; k < 8; ++k) cd[k] = dconst(cell.d[k], D);
            density = dlerp(dlerp(dlerp(cd[0], cd[4], fx), dlerp(cd[2], cd[6], fx), fy),
                            dlerp(dlerp(cd[1], cd[5], fx), dlerp(cd[3], cd[7], fx), fy), fz);
        }
        Dual<T> s[4];
        tf_eval_dual(in, fd, b, density, s);
        // blending, reference form acc + (1 - acc.a) * alpha * (rgb, 1)
        Dual<T> alpha;
        if (in.blend == BlendBeerLambert) {
            alpha = T(1) - dexp(T(-1) * (s[3] * step));
        } else {
            const Dual<T> xa = s[3] * step;
            alpha = xa.v < T(1) ? xa : dconst(T(1), D);
        }
        const Dual<T> w = (T(1) - acc[3]) * alpha;
        for (int c = 0; c < 3; ++c) acc[c] = acc[c] + w * s[c];
        acc[3] = acc[3] + w;
    }
    color = v4(acc[0].v, acc[1].v, acc[2].v, acc[3].v);
    term = i;
    for (int k = 0; k < D; ++k)
        for (int c = 0; c < 4; ++c) grad[4 * k + c] = acc[c].d[(size_t)k];
}

template <class T>
inline void render_forward_gradients(const Inputs<T>& in, const FwdDiffSettings& fd, T* color, int* term, T* grad,
                                     const int* frozen_term, const T* frozen_t0)
{
    const long long count = (long long)in.B * in.H * in.W;
#pragma omp parallel for schedule(guided)
    for (long long i = 0; i < count; ++i) {
        const int b = (int)(i / ((long long)in.W * in.H));
        const int y = (int)((i - (long long)b * in.W * in.H) / in.W);
        const int x = (int)(i - (long long)in.W * (y + (long long)in.H * b));
        V4<T> c; int n;
        march_forward_gradients(in, fd, x, y, b, frozen_term ? frozen_term[i] : -1,
                                frozen_t0 ? frozen_t0 + i : (const T*)nullptr, c, n, grad + (size_t)i * fd.D * 4);
        color[4 * i + 0] = c.x; color[4 * i + 1] = c.y; color[4 * i + 2] = c.z; color[4 * i + 3] = c.w;
        term[i] = n;
    }
}

// ---------------------------------------------------------------------------
// Kink census (test support).  Counts the samples of a ray (sample count n
// frozen) whose FP64 density lies within eps * (largest corner magnitude of
// its cell, the scale of FP32 interpolation error) of a point where the
// reference's per-sample derivative jumps by a finite amount:
//   identity TF: the clamp mask at density 0 and 1 (renderer_tf.cuh:62-73,103);
//   texture TF: texel edges, density*R - 0.5 integer (renderer_tf.cuh:160-171);
//   linear TF: control points (renderer_tf.cuh:294-313);
//   alpha blending: the clamp at absorption*step = 1 (renderer_blending.cuh:118-125);
//   nearest filtering: the rounding edges p = k + 0.5, where the density itself
//     jumps (renderer_interpolation.cuh:38-50); counted within eps_pos voxels;
//   trilinear / pre-shaded cell faces, p = k for k > 0 (renderer_interpolation.cuh:
//     218-225), where the position derivative jumps by a second difference of
//     the volume; counted within eps_face voxels when eps_face > 0 (p = 0 is not
//     a kink: truncation keeps both sides in cell 0).
// An FP32 evaluation may place such a sample on the other side of the kink
// than FP64 does; its adjoint then differs by the jump, which no FP32
// implementation of the reference's mathematics avoids.  Cell faces matter for
// the position and step gradients only (the step gradient weighs a sample's
// position derivative by its index), so they are counted on request.
// Pre-shaded volumes have no TF kinks; their alpha clamp uses the absorption
// channel.
// ---------------------------------------------------------------------------
template <class T>
inline int kink_samples(const Inputs<T>& in, int x, int y, int b, int n, T eps, T eps_pos, T eps_face)
{
    const Ray<T> r = setup_ray(in, x, y, b);
    const T* tf = in.tf + (size_t)bidx(in.tf_b, b) * in.R * in.C;
    const int vb = bidx(in.vol_b, b);
    int count = 0;
    for (int i = 0; i < n; ++i) {
        const T t = r.t0 + T(i) * r.step;
        const V3<T> p = sample_pos(r, t);
        // density and the magnitude its FP32 rounding error scales with; `exact`: every
        // precision computes the same density (a voxel value, or a cell whose corners are all
        // equal: a + t (b - a) = a), so comparisons of it with 0, 1 or a control point agree
        T d, scale;
        T absorption = T(0);
        bool exact = false;
        if (in.filter == FilterPreshaded) {
            TriCell<T> cells[4];
            absorption = preshaded_fetch(in, p, cells).w;
            d = scale = T(0);
        } else if (in.filter == FilterNearest) {
            d = fetch_density(in, b, p);
            scale = std::fabs(d);
            exact = true;
        } else {
            TriCell<T> c;
            trilinear_cell(in, vb, p, c);
            d = trilinear_value(c);
            scale = T(0);
            exact = true;
            for (int k = 0; k < 8; ++k) {
                scale = std::max(scale, std::fabs(c.d[k]));
                exact &= c.d[k] == c.d[0];
            }
        }
        const T band = eps * scale;
        bool kink = false;
        if (in.tf_mode == TFIdentity) {
            kink = !exact && (std::fabs(d) < band || std::fabs(d - T(1)) < band);
        } else if (in.tf_mode == TFTexture) {
            // d * R - 0.5 rounds differently per precision even for an exact d
            const T u = d * T(in.R) - T(0.5);
            kink = std::fabs(u - std::round(u)) < band * T(in.R);
        } else if (in.tf_mode == TFLinear) {
            for (int k = 0; k < in.R && !exact; ++k) kink |= std::fabs(d - tf[k * in.C + 4]) < band;
        }
        if (in.blend == BlendAlpha) {
            const T a = (in.filter == FilterPreshaded ? absorption : tf_eval(in, b, d).w) * r.step;
            kink |= std::fabs(a - T(1)) < eps;
        }
        const T q[3] = {p.x, p.y, p.z};
        if (in.filter == FilterNearest) {
            for (int k = 0; k < 3; ++k) kink |= std::fabs(q[k] - std::floor(q[k]) - T(0.5)) < eps_pos;
        } else if (eps_face > T(0)) {
            for (int k = 0; k < 3; ++k) {
                const T face = std::round(q[k]);
                kink |= face != T(0) && std::fabs(q[k] - face) < eps_face;
            }
        }
        count += kink ? 1 : 0;
    }
    return count;
}

template <class T>
inline void kink_census(const Inputs<T>& in, const int* term, T eps, T eps_pos, T eps_face, int* out)
{
    const long long count = (long long)in.B * in.H * in.W;
#pragma omp parallel for schedule(guided)
    for (long long i = 0; i < count; ++i) {
        const int b = (int)(i / ((long long)in.W * in.H));
        const int y = (int)((i - (long long)b * in.W * in.H) / in.W);
        const int x = (int)(i - (long long)in.W * (y + (long long)in.H * b));
        out[i] = kink_samples(in, x, y, b, term[i], eps, eps_pos, eps_face);
    }
}

// ---------------------------------------------------------------------------
// f4: auxiliary kernels.  Dense volumes are X x Y x Z with z fastest (C order).
// ---------------------------------------------------------------------------
template <class T> inline T sq(T v) { return v * v; }

// pre-shading (renderer/tf_utils_impl.cu:17-32): color[ch][v] = TF(density[v])
template <class T>
inline void preshade(const Inputs<T>& tf, const T* dens, int64_t n, T* color)
{
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < n; ++v) {
        const V4<T> c = tf_eval(tf, 0, dens[v]);
        color[v] = c.x; color[n + v] = c.y; color[2 * n + v] = c.z; color[3 * n + v] = c.w;
    }
}

// its adjoint (:96-121): adj_density[v] = d<adj_color[.][v], TF>/d density
template <class T>
inline void preshade_adjoint(const Inputs<T>& tf, const T* dens, int64_t n, const T* adj_color, T* adj_dens)
{
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < n; ++v) {
        const V4<T> a = v4(adj_color[v], adj_color[n + v], adj_color[2 * n + v], adj_color[3 * n + v]);
        adj_dens[v] = tf_adjoint(tf, 0, dens[v], a, (T*)nullptr);
    }
}

// Random streams of the GPU kernels: SplitMix64 (Steele, Lea and Flood 2014)
// and the 48-bit LCG of java.util.Random (multiplier 0x5DEECE66D, increment 11),
// whose 24 high bits make a float in [0, 1) (renderer/tf_utils_impl.cu:196-210).
inline uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
inline uint64_t lcg48_next(uint64_t s) { return (s * 0x5DEECE66Dull + 0xBull) & ((1ull << 48) - 1); }
inline double lcg48_unit(uint64_t s) { return double(uint32_t(s >> 24)) / 16777216.0; }
inline uint64_t best_fit_stream(uint64_t seed, uint64_t voxel) { return splitmix64(splitmix64(seed) + voxel) & ((1ull << 48) - 1); }

// cost of density d for one voxel (:239-262, :389-409)
template <class T>
inline T best_fit_cost(const Inputs<T>& tf, const T target[4], T d, T opacity_w, const T* nb, const bool* has, T neigh_w)
{
    const V4<T> c = tf_eval(tf, 0, d);
    const T dx = target[0] - c.x, dy = target[1] - c.y, dz = target[2] - c.z;
    T cost = (dx * dx + dy * dy + dz * dz) + opacity_w * std::log1p(std::fabs(target[3] - c.w));
    if (nb)
        for (int j = 0; j < 6; ++j)
            if (has[j]) cost += neigh_w * (d - nb[j]) * (d - nb[j]);
    return cost;
}

// best fit (:225-272, :351-420): color 4 x X x Y x Z, prev (or null) and out X x Y x Z;
// best_cost (or null) receives the winning cost
template <class T>
inline void best_fit(const Inputs<T>& tf, const T* color, int X, int Y, int Z, int num_samples, T opacity_w,
                     const T* prev, T neigh_w, uint64_t seed, T* out, T* best_cost)
{
    const int64_t n = (int64_t)X * Y * Z;
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < n; ++v) {
        const int x = int(v / ((int64_t)Y * Z)), y = int((v / Z) % Y), z = int(v % Z);
        const T target[4] = {color[v], color[n + v], color[2 * n + v], color[3 * n + v]};
        T nb[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
        bool has[6] = {x > 0, x < X - 1, y > 0, y < Y - 1, z > 0, z < Z - 1};
        const int64_t st[6] = {-(int64_t)Y * Z, (int64_t)Y * Z, -Z, Z, -1, 1};
        if (prev)
            for (int j = 0; j < 6; ++j)
                if (has[j]) nb[j] = prev[v + st[j]];
        uint64_t s = best_fit_stream(seed, (uint64_t)v);
        T best = T(3.402823466e38), bestD = T(0);
        for (int k = 0; k < num_samples; ++k) {
            s = lcg48_next(s);
            const T d = T(lcg48_unit(s));
            const T cost = best_fit_cost(tf, target, d, opacity_w, prev ? nb : nullptr, has, neigh_w);
            if (cost < best) {
                best = cost;
                bestD = d;
            }
        }
        out[v] = bestD;
        if (best_cost) best_cost[v] = best;
    }
}

// trilinear fetch of a dense volume at voxel position p (renderer_interpolation.cuh:145-177)
template <class T> inline T fetch_dense_trilinear(const T* vol, int X, int Y, int Z, T px, T py, T pz)
{
    Inputs<T> in{};
    in.volume = vol; in.vol_b = 1; in.X = X; in.Y = Y; in.Z = Z; in.filter = FilterTrilinear;
    TriCell<T> c;
    trilinear_cell(in, 0, v3(px, py, pz), c);
    return trilinear_value(c);
}

// importance sampling (renderer/losses_impl.cu:254-292): maxima of |grad density|,
// opacity and |grad opacity| over the voxels (central differences, clamped at the border)
template <class T>
inline void importance_maxima(const Inputs<T>& tf, const T* dens, int X, int Y, int Z, T out[3])
{
    out[0] = out[1] = out[2] = T(0);
    auto at = [&](int x, int y, int z) { return dens[((int64_t)x * Y + y) * Z + z]; };
    auto op = [&](T d) { return tf_eval(tf, 0, d).w; };
    for (int x = 0; x < X; ++x)
        for (int y = 0; y < Y; ++y)
            for (int z = 0; z < Z; ++z) {
                const T c = at(x, y, z);
                const T xm = at(std::max(0, x - 1), y, z), xp = at(std::min(X - 1, x + 1), y, z);
                const T ym = at(x, std::max(0, y - 1), z), yp = at(x, std::min(Y - 1, y + 1), z);
                const T zm = at(x, y, std::max(0, z - 1)), zp = at(x, y, std::min(Z - 1, z + 1));
                const T gd = std::sqrt(sq((xp - xm) / 2) + sq((yp - ym) / 2) + sq((zp - zm) / 2));
                const T go = std::sqrt(sq((op(xp) - op(xm)) / 2) + sq((op(yp) - op(ym)) / 2) + sq((op(zp) - op(zm)) / 2));
                out[0] = std::max(out[0], gd);
                out[1] = std::max(out[1], op(c));
                out[2] = std::max(out[2], go);
            }
}

// the decision for sample i (:298-371) with the GPU's uniform numbers in (0, 1]
// (two 24-bit fields of splitmix64(splitmix64(seed) + i)); weights are the
// cumulative normalised thresholds (losses.cpp:188-200, in float as there).
// margin[i] = |value / max - rndReject| for the rejection branches, +inf for
// the uniform branch: decisions with a small margin may flip under rounding.
template <class T>
inline void sample_importance(const Inputs<T>& tf, const T* dens, int X, int Y, int Z, const T* loc, int64_t N,
                              const float w[3], const T maxima[3], uint32_t seed, uint8_t* out, T* margin)
{
    const uint64_t sm = splitmix64((uint64_t)seed);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        const T px = loc[3 * i] * T(X - 1), py = loc[3 * i + 1] * T(Y - 1), pz = loc[3 * i + 2] * T(Z - 1);
        const uint64_t h = splitmix64(sm + (uint64_t)i);
        const float rndSample = float(uint32_t(h >> 40) + 1u) / 16777216.0f;
        const T rndReject = T(float(uint32_t((h >> 16) & 0xFFFFFFull) + 1u) / 16777216.0f);
        auto f = [&](T dx, T dy, T dz) { return fetch_dense_trilinear(dens, X, Y, Z, px + dx, py + dy, pz + dz); };
        auto op = [&](T d) { return tf_eval(tf, 0, d).w; };
        T value;
        if (rndSample <= w[0]) {
            out[i] = 1;
            if (margin) margin[i] = std::numeric_limits<T>::infinity();
            continue;
        } else if (rndSample <= w[1]) {
            value = std::sqrt(sq((f(1, 0, 0) - f(-1, 0, 0)) / 2) + sq((f(0, 1, 0) - f(0, -1, 0)) / 2) +
                              sq((f(0, 0, 1) - f(0, 0, -1)) / 2)) / maxima[0];
        } else if (rndSample <= w[2]) {
            value = op(f(0, 0, 0)) / maxima[1];
        } else {
            value = std::sqrt(sq((op(f(1, 0, 0)) - op(f(-1, 0, 0))) / 2) + sq((op(f(0, 1, 0)) - op(f(0, -1, 0))) / 2) +
                              sq((op(f(0, 0, 1)) - op(f(0, 0, -1))) / 2)) / maxima[2];
        }
        out[i] = value >= rndReject ? 1 : 0;
        if (margin) margin[i] = std::fabs(value - rndReject);
    }
}

}  // namespace oracle
