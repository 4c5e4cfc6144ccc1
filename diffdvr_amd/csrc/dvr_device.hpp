// Device building blocks of the ray marcher for gfx950 (wave64).
//
// Semantics follow DiffDVR (paths relative to that repository):
//   camera modes      renderer/renderer_camera.cuh:12-61
//   slab test         renderer/renderer_utils.cuh:60-74 (tmin clamped to 0, renderer_kernels.cuh:48)
//   interpolation     renderer/renderer_interpolation.cuh:33-389
//   transfer function renderer/renderer_tf.cuh:59-648
//   blending          renderer/renderer_blending.cuh:36-165
// The implementation is this project's own: float march (expf in the plain
// kernels; the Beer-Lambert keep and the FAST Gaussian weights on the hardware
// base-2 exponential), double per-ray setup, 32-bit offsets.
#pragma once
#include <hip/hip_runtime.h>
#include "dvr_params.hpp"

namespace dvr {

struct f3 { float x, y, z; };
struct f4 { float x, y, z, w; };

__device__ __forceinline__ f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f4 mk4(float x, float y, float z, float w) { return f4{x, y, z, w}; }
__device__ __forceinline__ f3 add3(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3 sub3(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 mul3(f3 a, f3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ f3 div3(f3 a, f3 b) { return mk3(a.x / b.x, a.y / b.y, a.z / b.z); }
__device__ __forceinline__ f3 scl3(f3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross3(f3 a, f3 b)
{
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ f3 normalize3(f3 v) { return scl3(v, rsqrtf(dot3(v, v))); }
__device__ __forceinline__ float lerpf_(float a, float b, float t) { return a + t * (b - a); }
// ternary min/max as in the reference (helper_math.cuh:86-104)
__device__ __forceinline__ float tmin_(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float tmax_(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float fclamp(float f, float a, float b) { return fmaxf(a, fminf(f, b)); }
__device__ __forceinline__ int iclamp(int f, int a, int b) { return max(a, min(f, b)); }

// Adjoints of normalize and cross as the reference forms them
// (renderer_adjoint.cuh:28-65).
// adjNormalize (renderer_adjoint.cuh:28-45)
__device__ __forceinline__ f3 adj_normalize(f3 in, f3 adj_out)
{
    const float len = sqrtf(dot3(in, in));
    f3 adj_in = scl3(adj_out, 1.0f / len);
    const float adj_len = dot3(adj_out, scl3(in, -1.0f / (len * len)));
    return add3(adj_in, scl3(in, adj_len / len));
}

// adjCross (renderer_adjoint.cuh:47-65)
__device__ __forceinline__ void adj_cross(f3 a, f3 b, f3 c, f3& aa, f3& ab)
{
    aa.y += b.z * c.x; ab.z += a.y * c.x; aa.z -= b.y * c.x; ab.y -= a.z * c.x;
    aa.z += b.x * c.y; ab.x += a.z * c.y; aa.x -= b.z * c.y; ab.z -= a.x * c.y;
    aa.x += b.y * c.z; ab.y += a.x * c.z; aa.y -= b.x * c.z; ab.x -= a.y * c.z;
}

// ---------------------------------------------------------------------------
// Global memory through a uniform 64-bit base and a 32-bit byte offset per
// lane (offsets < 2^32 bytes: slices of at most 2^30 floats).  The compiler
// then emits global_load / global_store / global_atomic in their SGPR-base
// form instead of flat instructions: a flat access, which it must use for a
// pointer it cannot prove global (one read from a table), also counts on the
// LDS wait counter, so every wait for an LDS read would wait for the voxel
// gathers in flight as well.
// ---------------------------------------------------------------------------
typedef __attribute__((address_space(1))) float gfloat;
typedef __attribute__((address_space(1))) int gint;
typedef __attribute__((address_space(1))) char gchar;

__device__ __forceinline__ float gld(const float* base, unsigned idx)
{
    return *(const gfloat*)((const gchar*)base + idx * 4u);
}
__device__ __forceinline__ void gst(float* base, unsigned idx, float v) { *(gfloat*)((gchar*)base + idx * 4u) = v; }
__device__ __forceinline__ void gsti(int* base, unsigned idx, int v) { *(gint*)((gchar*)base + idx * 4u) = v; }
__device__ __forceinline__ void gatomic_add(float* base, unsigned idx, float v)
{
    __hip_atomic_fetch_add((gfloat*)((gchar*)base + idx * 4u), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// Wave-level helpers (wave64)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Workgroup index -> logical tile index so that each XCD (workgroups are
// dispatched round-robin over the 8 XCDs) owns one contiguous range of tiles:
// neighbouring tiles then share that XCD's L2.
__device__ __forceinline__ int xcd_remap(int bid, int nblocks)
{
    const int xcd = bid % kNumXcd, slot = bid / kNumXcd;
    const int base = nblocks / kNumXcd, rem = nblocks % kNumXcd;
    return xcd * base + min(xcd, rem) + slot;
}

// ---------------------------------------------------------------------------
// Per-ray setup (renderer_kernels.cuh:31-48).  Once per ray, the camera ray
// and the slab test are evaluated in double and rounded to float: the
// per-sample march then starts from correctly rounded ray parameters (a few
// ulp of direction error shift distant samples by ~1e-5 voxel, which sharp
// transfer functions amplify).  The march itself is float.
// ---------------------------------------------------------------------------
struct Ray {
    f3 start, dir;
    f3 bmin, voxel, invVoxel;
    f3 p0, dp;  // sample position in voxel units: p(t) = p0 + t * dp
    float step, t0, t1;
};

struct d3 {
    double x, y, z;
};
__device__ __forceinline__ d3 dk3(double x, double y, double z) { return d3{x, y, z}; }
__device__ __forceinline__ d3 dcross(d3 a, d3 b)
{
    return dk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

__device__ __forceinline__ d3 matmul_proj_d(const double m[16], double x, double y, double z)
{
    const double w0 = m[0] * x + m[1] * y + m[2] * z + m[3];
    const double w1 = m[4] * x + m[5] * y + m[6] * z + m[7];
    const double w2 = m[8] * x + m[9] * y + m[10] * z + m[11];
    const double w3 = m[12] * x + m[13] * y + m[14] * z + m[15];
    return dk3(w0 / w3, w1 / w3, w2 / w3);
}

__device__ __forceinline__ d3 dnormalize(d3 v)
{
    const double inv = 1.0 / sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    return dk3(v.x * inv, v.y * inv, v.z * inv);
}

// InverseViewMatrix camera: unproject NDC (fx, fy) at depths 0 and 0.9
// (renderer_camera.cuh:28-44, renderer_utils.cuh:53-58).
__device__ __forceinline__ void camera_inverse_view(const double m[16], int x, int y, int W, int H, d3& start, d3& dir)
{
    const double fx = 2.0 * (x + 0.5) / double(W) - 1.0;
    const double fy = 2.0 * (y + 0.5) / double(H) - 1.0;
    start = matmul_proj_d(m, fx, fy, 0.0);
    const d3 end = matmul_proj_d(m, fx, fy, 0.9);
    dir = dnormalize(dk3(end.x - start.x, end.y - start.y, end.z - start.z));
}

// ReferenceFrame camera: viewport rows eye, right, up; front = up x right
// (renderer_camera.cuh:46-61).
__device__ __forceinline__ void camera_reference_frame(d3 eye, d3 right, d3 up, double tanX, double tanY, int x, int y,
                                                       int W, int H, d3& start, d3& dir)
{
    const d3 front = dcross(up, right);
    const double fx = (2.0 * (x + 0.5) / double(W) - 1.0) * tanX;
    const double fy = (2.0 * (y + 0.5) / double(H) - 1.0) * tanY;
    start = eye;
    dir = dnormalize(dk3(front.x + fx * right.x + fy * up.x, front.y + fx * right.y + fy * up.y,
                         front.z + fx * right.z + fy * up.z));
}

// Camera from the standard parameter block (renderer_camera.cuh:12-61).
__device__ __forceinline__ void camera_ray(const RenderParams& P, int x, int y, int b, d3& start, d3& dir)
{
    if (P.camMode == 0) {  // RayStartDir
        const float* s = P.rs + b * P.rsSb + y * P.rsSy + x * P.rsSx;
        const float* d = P.rd + b * P.rdSb + y * P.rdSy + x * P.rdSx;
        start = dk3(s[0], s[P.rsSc], s[2 * P.rsSc]);
        dir = dk3(d[0], d[P.rdSc], d[2 * P.rdSc]);
    } else if (P.camMode == 1) {  // InverseViewMatrix
        double m[16];
        if (P.cm) {
            const float* c = P.cm + b * P.cmSb;
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int k = 0; k < 4; ++k) m[4 * r + k] = c[r * P.cmSr + k * P.cmSc];
        } else {
#pragma unroll
            for (int k = 0; k < 16; ++k) m[k] = P.cmU[k];
        }
        camera_inverse_view(m, x, y, P.W, P.H, start, dir);
    } else {  // ReferenceFrame
        const float* c = P.cm + b * P.cmSb;
        const d3 eye = dk3(c[0], c[P.cmSc], c[2 * P.cmSc]);
        const d3 right = dk3(c[P.cmSr], c[P.cmSr + P.cmSc], c[P.cmSr + 2 * P.cmSc]);
        const d3 up = dk3(c[2 * P.cmSr], c[2 * P.cmSr + P.cmSc], c[2 * P.cmSr + 2 * P.cmSc]);
        camera_reference_frame(eye, right, up, P.tanX, P.tanY, x, y, P.W, P.H, start, dir);
    }
}

// Slab test with the box given as min/size (renderer_utils.cuh:60-74), in double.
__device__ __forceinline__ void slab(d3 start, d3 dir, d3 bmin, d3 bsize, double& t0, double& t1)
{
    const double ix = 1.0 / dir.x, iy = 1.0 / dir.y, iz = 1.0 / dir.z;
    const double a1 = (bmin.x - start.x) * ix, a2 = (bmin.x + bsize.x - start.x) * ix;
    const double a3 = (bmin.y - start.y) * iy, a4 = (bmin.y + bsize.y - start.y) * iy;
    const double a5 = (bmin.z - start.z) * iz, a6 = (bmin.z + bsize.z - start.z) * iz;
    auto mn = [](double a, double b) { return a < b ? a : b; };
    auto mx = [](double a, double b) { return a > b ? a : b; };
    t0 = mx(mx(mn(a1, a2), mn(a3, a4)), mn(a5, a6));
    t1 = mn(mn(mx(a1, a2), mx(a3, a4)), mx(a5, a6));
}

// Ray of a camera ray through the box [bmin, bmin + bsize] of an X x Y x Z
// volume (renderer_kernels.cuh:31-48), rounded from double.
__device__ __forceinline__ Ray make_ray(d3 start, d3 dir, d3 bmin, d3 bsize, int X, int Y, int Z, float step)
{
    Ray r;
    r.bmin = mk3(float(bmin.x), float(bmin.y), float(bmin.z));
    r.voxel = div3(mk3(float(bsize.x), float(bsize.y), float(bsize.z)), mk3(float(X - 1), float(Y - 1), float(Z - 1)));
    r.invVoxel = mk3(1.0f / r.voxel.x, 1.0f / r.voxel.y, 1.0f / r.voxel.z);
    r.start = mk3(float(start.x), float(start.y), float(start.z));
    r.dir = mk3(float(dir.x), float(dir.y), float(dir.z));
    r.step = step;
    {
        const double vx = bsize.x / double(X - 1), vy = bsize.y / double(Y - 1), vz = bsize.z / double(Z - 1);
        r.p0 = mk3(float((start.x - bmin.x) / vx), float((start.y - bmin.y) / vy), float((start.z - bmin.z) / vz));
        r.dp = mk3(float(dir.x / vx), float(dir.y / vy), float(dir.z / vz));
    }
    double t0, t1;
    slab(start, dir, bmin, bsize, t0, t1);
    t0 = t0 > 0.0 ? t0 : 0.0;
    r.t0 = float(t0);
    r.t1 = float(t1);
    // Degenerate rays (zero direction, NaN slab bounds) and non-positive or
    // non-finite steps would never leave the march loop: they sample nothing.
    if (!(r.t1 >= r.t0) || !isfinite(r.t0) || !isfinite(r.t1) || !(r.step > 0.0f) || !isfinite(r.step)) {
        r.t0 = 0.0f;
        r.t1 = -1.0f;
        if (!(r.step > 0.0f) || !isfinite(r.step)) r.step = 1.0f;
    }
    return r;
}

__device__ __forceinline__ Ray setup_ray(const RenderParams& P, int x, int y, int b)
{
    d3 bmin, bsize;
    if (P.bmin) {
        const float* p = P.bmin + b * P.bminSb;
        bmin = dk3(p[0], p[P.bminSc], p[2 * P.bminSc]);
    } else {
        bmin = dk3(P.bminU[0], P.bminU[1], P.bminU[2]);
    }
    if (P.bsize) {
        const float* p = P.bsize + b * P.bsizeSb;
        bsize = dk3(p[0], p[P.bsizeSc], p[2 * P.bsizeSc]);
    } else {
        bsize = dk3(P.bsizeU[0], P.bsizeU[1], P.bsizeU[2]);
    }
    d3 start, dir;
    camera_ray(P, x, y, b, start, dir);
    const float step = P.st ? P.st[b * P.stSb + y * P.stSy + x * P.stSx] : P.stU;
    return make_ray(start, dir, bmin, bsize, P.X, P.Y, P.Z, step);
}

// volume-space sample position of parameter t: ((start + t dir) - bmin) / voxel,
// as one FMA per axis from per-ray constants rounded from double
__device__ __forceinline__ f3 sample_pos(const Ray& r, float t)
{
    return mk3(fmaf(t, r.dp.x, r.p0.x), fmaf(t, r.dp.y, r.p0.y), fmaf(t, r.dp.z, r.p0.z));
}

// ---------------------------------------------------------------------------
// Volume access (one batch slice)
// ---------------------------------------------------------------------------
struct Vol {
    const float* p;   // batch slice base
    int X, Y, Z;
    int sx, sy, sz;
};

__device__ __forceinline__ Vol make_vol(const RenderParams& P, int b)
{
    return Vol{P.vol + b * P.vsb, P.X, P.Y, P.Z, P.vsx, P.vsy, P.vsz};
}

// Trilinear cell cache: the 8 corners of the current cell are reloaded only
// when the (truncated) cell index changes (renderer_interpolation.cuh:218-225).
struct Cell {
    int ix, iy, iz;
    float d[8];  // bit order (x<<2)|(y<<1)|z
};

__device__ __forceinline__ void cell_invalidate(Cell& c) { c.ix = c.iy = c.iz = -0x40000000; }

// Corner offsets of cell (ix, iy, iz) with the reference's per-corner clamp
// (renderer_interpolation.cuh:150-157): the low corner is clamp(i, 0, n-1)
// and the high corner clamp(i+1, 0, n-1), so the high corner is one stride
// further exactly when 0 <= i < n-1, else it coincides with the low corner.
// One multiply per axis instead of two (v_mul_lo_u32 is quarter rate).
struct CellOffsets {
    unsigned base, dx, dy, dz;
};

__device__ __forceinline__ CellOffsets cell_offsets(int ix, int iy, int iz, int X, int Y, int Z, int sx, int sy, int sz)
{
    CellOffsets o;
    o.base = unsigned(iclamp(ix, 0, X - 1)) * unsigned(sx) + unsigned(iclamp(iy, 0, Y - 1)) * unsigned(sy) +
             unsigned(iclamp(iz, 0, Z - 1)) * unsigned(sz);
    o.dx = unsigned(ix) < unsigned(X - 1) ? unsigned(sx) : 0u;
    o.dy = unsigned(iy) < unsigned(Y - 1) ? unsigned(sy) : 0u;
    o.dz = unsigned(iz) < unsigned(Z - 1) ? unsigned(sz) : 0u;
    return o;
}

__device__ __forceinline__ void cell_load(const Vol& V, int ix, int iy, int iz, float d[8])
{
    const CellOffsets o = cell_offsets(ix, iy, iz, V.X, V.Y, V.Z, V.sx, V.sy, V.sz);
    const unsigned b = o.base, xy = o.dx + o.dy;
    d[0] = gld(V.p, b);
    d[1] = gld(V.p, b + o.dz);
    d[2] = gld(V.p, b + o.dy);
    d[3] = gld(V.p, b + o.dy + o.dz);
    d[4] = gld(V.p, b + o.dx);
    d[5] = gld(V.p, b + o.dx + o.dz);
    d[6] = gld(V.p, b + xy);
    d[7] = gld(V.p, b + xy + o.dz);
}

__device__ __forceinline__ float tri_lerp(const float d[8], float fx, float fy, float fz)
{
    return lerpf_(lerpf_(lerpf_(d[0], d[4], fx), lerpf_(d[2], d[6], fx), fy),
                  lerpf_(lerpf_(d[1], d[5], fx), lerpf_(d[3], d[7], fx), fy), fz);
}

// Nearest neighbour index (renderer_interpolation.cuh:38-50)
__device__ __forceinline__ int nearest_offset(const Vol& V, f3 p)
{
    const float eps = 1e-7f;
    const float cx = fclamp(p.x, eps, float(V.X) - (1.0f + eps));
    const float cy = fclamp(p.y, eps, float(V.Y) - (1.0f + eps));
    const float cz = fclamp(p.z, eps, float(V.Z) - (1.0f + eps));
    return int(roundf(cx)) * V.sx + int(roundf(cy)) * V.sy + int(roundf(cz)) * V.sz;
}


template <int FILTER>
__device__ __forceinline__ float fetch(const Vol& V, f3 p, Cell& c)
{
    if constexpr (FILTER == 0) {
        return gld(V.p, unsigned(nearest_offset(V, p)));
    } else {
        const int ix = int(p.x), iy = int(p.y), iz = int(p.z);
        if (ix != c.ix || iy != c.iy || iz != c.iz) {
            cell_load(V, ix, iy, iz, c.d);
            c.ix = ix; c.iy = iy; c.iz = iz;
        }
        return tri_lerp(c.d, p.x - float(ix), p.y - float(iy), p.z - float(iz));
    }
}

// ---------------------------------------------------------------------------
// Fetch pipeline: the voxels of the NEXT sample are requested before the
// current sample is shaded and blended, so the L1/L2 round trip overlaps a
// sample's worth of arithmetic instead of stalling the wave (a wave of 8x8
// rays enters a new cell on almost every step, so a plain cell cache issues
// and waits for loads each sample).  Positions handed to request() must be
// the ones later passed to value(): callers carry them from one iteration
// to the next, and call value(p) before request(V, pn): the next cell's
// loads land in the registers of the current cell, whose values are no longer
// needed once p is interpolated (one set of 8 corner registers, not two).
//   start(V, p)   load the cell of the first position
//   value(p)      interpolate at p (p lies in the current cell)
//   request(V, pn) issue the loads of the cell of pn if it differs; pn's cell
//                 is current from here on (its first use waits for the loads)
//   advance()     nothing left to do (kept for the nearest filter's pipe)
// ---------------------------------------------------------------------------
template <int FILTER>
struct FetchPipe;

template <>
struct FetchPipe<1> {
    int ix, iy, iz;
    float d[8];
    __device__ __forceinline__ void start(const Vol& V, f3 p)
    {
        ix = int(p.x); iy = int(p.y); iz = int(p.z);
        cell_load(V, ix, iy, iz, d);
    }
    __device__ __forceinline__ float value(f3 p) const
    {
        return tri_lerp(d, p.x - float(ix), p.y - float(iy), p.z - float(iz));
    }
    __device__ __forceinline__ void request(const Vol& V, f3 pn)
    {
        const int jx = int(pn.x), jy = int(pn.y), jz = int(pn.z);
        if ((jx != ix) | (jy != iy) | (jz != iz)) {
            cell_load(V, jx, jy, jz, d);
            ix = jx; iy = jy; iz = jz;
        }
    }
    __device__ __forceinline__ void advance() {}
};

template <>
struct FetchPipe<0> {
    float v, nv;
    __device__ __forceinline__ void start(const Vol& V, f3 p) { v = gld(V.p, unsigned(nearest_offset(V, p))); }
    __device__ __forceinline__ float value(f3) const { return v; }
    __device__ __forceinline__ void request(const Vol& V, f3 pn) { nv = gld(V.p, unsigned(nearest_offset(V, pn))); }
    __device__ __forceinline__ void advance() { v = nv; }
};

// ---------------------------------------------------------------------------
// Transfer functions.  `tf` points at row 0 of this batch's R x C table with
// row stride sr and column stride sc (global or LDS memory).
// ---------------------------------------------------------------------------
struct Tf {
    const float* p;
    int R, C, sr, sc;
};

// Columns of a TF table per mode (identity, texture, linear, Gaussian); LDS
// copies are staged densely, so kernels that read them pass this constant as
// row stride and the row addressing needs no multiply.
template <int TFM>
constexpr int tf_cols()
{
    return TFM == 0 ? 2 : (TFM == 1 ? 4 : (TFM == 2 ? 5 : 6));
}
// Row stride of a TF staged in LDS by the batched kernel: the linear TF's rows
// are padded to 8 floats so a row's RGBA is one 16-byte LDS read.
template <int TFM>
constexpr int tf_lds_stride()
{
    return TFM == 2 ? 8 : tf_cols<TFM>();
}

// Interval of a piecewise-linear TF that holds a density: rows i and i+1, their
// control-point positions, and (FAST) the reciprocal of the interval width.
// FAST instances (the tiled, batched and forward-mode kernels) multiply by
// v_rcp_f32's reciprocal (1 ulp) instead of dividing, for the linear and the
// Gaussian TFs; the plain kernels keep the reference's divisions
// (renderer_tf.cuh:309,337-357,505-519,580-600).
struct TfLoc {
    int i;
    float p0, p1, inv;
};
// The interval that starts at row i (a density's interval once its row is known).
template <bool FAST>
__device__ __forceinline__ TfLoc tf_linear_row(const Tf& T, int i)
{
    TfLoc L;
    L.i = i;
    L.p0 = T.p[i * T.sr + 4 * T.sc];
    L.p1 = T.p[(i + 1) * T.sr + 4 * T.sc];
    L.inv = FAST ? __builtin_amdgcn_rcpf(L.p1 - L.p0) : 0.0f;
    return L;
}
template <bool FAST>
__device__ __forceinline__ TfLoc tf_linear_locate(const Tf& T, float density)
{
    int i;
    for (i = 0; i < T.R - 2; ++i)
        if (T.p[(i + 1) * T.sr + 4 * T.sc] > density) break;
    return tf_linear_row<FAST>(T, i);
}

// One Gaussian component's weight exp(-(d - mu)^2 / (2 s^2)); FAST multiplies by
// v_rcp_f32's reciprocal instead of dividing and evaluates 2^(-z log2 e) on the
// hardware base-2 exponential, z = (d - mu)^2 / (2 s^2) (relative error about
// z * 1e-7, below 2e-6 for the weights above 1e-8), in one expression for the
// value, the absorption and the adjoint, so the three agree bitwise.
template <bool FAST>
__device__ __forceinline__ float gauss_n(float density, float mu, float s)
{
    if constexpr (FAST)
        return __builtin_amdgcn_exp2f(-(density - mu) * (density - mu) *
                                      (1.44269504088896341f * __builtin_amdgcn_rcpf(2.0f * s * s)));
    else return expf(-(density - mu) * (density - mu) / (2.0f * s * s));
}

// A sample's TF evaluation shared by its value and its adjoint in the reverse
// walk: the linear TF's located interval, or the first kGaussCache Gaussian
// weights (each an exponential, otherwise evaluated twice per sample).
constexpr int kGaussCache = 8;
struct TfPre {
    TfLoc L;
    float gn[kGaussCache];
};
template <int TFM, bool FAST>
__device__ __forceinline__ void tf_prepare(const Tf& T, float density, TfPre& pre)
{
    if constexpr (TFM == 2) pre.L = tf_linear_locate<FAST>(T, density);
    if constexpr (TFM == 3) {
#pragma unroll
        for (int i = 0; i < kGaussCache; ++i)
            pre.gn[i] = i < T.R ? gauss_n<FAST>(density, T.p[i * T.sr + 4 * T.sc], T.p[i * T.sr + 5 * T.sc]) : 0.0f;
    }
}

template <int TFM, bool FAST = false>
__device__ __forceinline__ f4 tf_eval(const Tf& T, float density, const TfPre* pre = nullptr)
{
    if constexpr (TFM == 0) {  // identity (renderer_tf.cuh:62-73)
        const float d = fclamp(density, 0.0f, 1.0f);
        const float sa = T.p[0], sc = T.p[T.sc];
        return mk4(d * sc, d * sc, d * sc, d * sa);
    } else if constexpr (TFM == 1) {  // texture (:160-171)
        const float d = density * float(T.R) - 0.5f;
        const int di = int(floorf(d));
        const float df = d - float(di);
        const float* a = T.p + iclamp(di, 0, T.R - 1) * T.sr;
        const float* c = T.p + iclamp(di + 1, 0, T.R - 1) * T.sr;
        return mk4(lerpf_(a[0], c[0], df), lerpf_(a[T.sc], c[T.sc], df), lerpf_(a[2 * T.sc], c[2 * T.sc], df),
                   lerpf_(a[3 * T.sc], c[3 * T.sc], df));
    } else if constexpr (TFM == 2) {  // piecewise linear (:294-313)
        const TfLoc L = pre ? pre->L : tf_linear_locate<FAST>(T, density);
        const float* a = T.p + L.i * T.sr;
        const float* c = a + T.sr;
        const float p0 = L.p0, p1 = L.p1;
        const float d2 = fclamp(density, p0, p1);
        const float f = FAST ? (d2 - p0) * L.inv : (d2 - p0) / (p1 - p0);
        return mk4(lerpf_(a[0], c[0], f), lerpf_(a[T.sc], c[T.sc], f), lerpf_(a[2 * T.sc], c[2 * T.sc], f),
                   lerpf_(a[3 * T.sc], c[3 * T.sc], f));
    } else {  // sum of Gaussians, column 5 = sigma (:491-519)
        f4 o = mk4(0.f, 0.f, 0.f, 0.f);
        auto comp = [&](int i, float n) __attribute__((always_inline)) {
            const float* r = T.p + i * T.sr;
            o.x += r[0] * n; o.y += r[T.sc] * n; o.z += r[2 * T.sc] * n; o.w += r[3 * T.sc] * n;
        };
        int i0 = 0;
        if (pre) {
#pragma unroll
            for (int i = 0; i < kGaussCache; ++i)
                if (i < T.R) comp(i, pre->gn[i]);
            i0 = kGaussCache;
        }
        for (int i = i0; i < T.R; ++i) {
            const float* r = T.p + i * T.sr;
            comp(i, gauss_n<FAST>(density, r[4 * T.sc], r[5 * T.sc]));
        }
        return o;
    }
}

// Absorption channel of the TF only (the transmittance recompute needs no colour).
// A linear TF's interval is located here unless the caller passes it in `loc`.
template <int TFM, bool FAST = false>
__device__ __forceinline__ float tf_absorption(const Tf& T, float density, const TfLoc* loc = nullptr)
{
    if constexpr (TFM == 0) {
        return fclamp(density, 0.0f, 1.0f) * T.p[0];
    } else if constexpr (TFM == 1) {
        const float d = density * float(T.R) - 0.5f;
        const int di = int(floorf(d));
        const float df = d - float(di);
        const float a = T.p[iclamp(di, 0, T.R - 1) * T.sr + 3 * T.sc];
        const float c = T.p[iclamp(di + 1, 0, T.R - 1) * T.sr + 3 * T.sc];
        return lerpf_(a, c, df);
    } else if constexpr (TFM == 2) {
        const TfLoc L = loc ? *loc : tf_linear_locate<FAST>(T, density);
        const float* a = T.p + L.i * T.sr;
        const float* c = a + T.sr;
        const float d2 = fclamp(density, L.p0, L.p1);
        const float f = FAST ? (d2 - L.p0) * L.inv : (d2 - L.p0) / (L.p1 - L.p0);
        return lerpf_(a[3 * T.sc], c[3 * T.sc], f);
    } else {
        float o = 0.0f;
        for (int i = 0; i < T.R; ++i) {
            const float* r = T.p + i * T.sr;
            const float mu = r[4 * T.sc], s = r[5 * T.sc];
            o += r[3 * T.sc] * gauss_n<FAST>(density, mu, s);
        }
        return o;
    }
}

// ---------------------------------------------------------------------------
// Blending (renderer_blending.cuh:54-61, 118-125)
// ---------------------------------------------------------------------------
template <int BLEND>
__device__ __forceinline__ float blend_alpha(float absorption, float step)
{
    if constexpr (BLEND == 0) return 1.0f - expf(-absorption * step);
    else return tmin_(1.0f, absorption * step);
}

template <int BLEND>
__device__ __forceinline__ void blend(f4& acc, const f4& s, float step)
{
    const float a = blend_alpha<BLEND>(s.w, step);
    const float tr = 1.0f - acc.w;
    acc.x = acc.x + tr * s.x * a;
    acc.y = acc.y + tr * s.y * a;
    acc.z = acc.z + tr * s.z * a;
    acc.w = acc.w + tr * a;
}

// ---------------------------------------------------------------------------
// Transmittance-tracking form used by the tiled kernels.  Mathematically the
// same blend as above, evaluated so that FP32 keeps its relative precision:
//   alpha = 1 - exp(-x) is formed with expm1 for small x (no cancellation),
//   keep  = 1 - alpha   is exp(-x) itself,
//   T     = 1 - A       is carried multiplicatively (T *= keep), so factors
//                       (1 - A) stay accurate as A approaches 1, while A is
//                       summed (A += T alpha) and stays accurate while small.
// ---------------------------------------------------------------------------
struct Acc {
    float r, g, b, A, T;
};

// Beer-Lambert: keep = exp(-x), alpha = 1 - exp(-x).  For x < 0.5, alpha is
// its Taylor series to degree 8 (truncation < 2e-8 relative), free of the
// cancellation of 1 - keep; for larger x, alpha = 1 - keep > 0.39 is well
// conditioned.  Both forms are evaluated and selected without a branch, so
// the fetches in flight around a sample are not split by divergent control flow.
// keep is 2^(-x log2 e) on the hardware base-2 exponential (v_exp_f32, ~1 ulp):
// rounding the product adds |x| * 6e-8 relative error, so a ray's transmittance
// carries at most (optical depth) * 6e-8, below 7e-7 before T reaches 1e-5.
template <int BLEND>
__device__ __forceinline__ void alpha_keep(float absorption, float step, float& alpha, float& keep)
{
    if constexpr (BLEND == 0) {
        const float x = absorption * step;
        keep = __builtin_amdgcn_exp2f(x * -1.44269504088896341f);
        const float xs = fminf(x, 0.5f);
        float p = -1.0f / 40320.0f;
        p = fmaf(p, xs, 1.0f / 5040.0f);
        p = fmaf(p, xs, -1.0f / 720.0f);
        p = fmaf(p, xs, 1.0f / 120.0f);
        p = fmaf(p, xs, -1.0f / 24.0f);
        p = fmaf(p, xs, 1.0f / 6.0f);
        p = fmaf(p, xs, -0.5f);
        p = fmaf(p, xs, 1.0f);
        alpha = x < 0.5f ? p * xs : 1.0f - keep;
    } else {
        alpha = tmin_(1.0f, absorption * step);
        keep = 1.0f - alpha;
    }
}

template <int BLEND>
__device__ __forceinline__ void blend_t(Acc& acc, const f4& s, float step)
{
    float alpha, keep;
    alpha_keep<BLEND>(s.w, step, alpha, keep);
    const float w = acc.T * alpha;
    acc.r += w * s.x;
    acc.g += w * s.y;
    acc.b += w * s.z;
    acc.A += w;
    acc.T *= keep;
}

// Blend adjoint given the transmittance before the sample (t_in = 1 - a_in);
// same derivative as blend_adjoint.  The alpha-mode clamp at 1-1e-5 follows
// renderer_blending.cuh:136,158-159.
// blend_adjoint_ak: the same with the Beer-Lambert (alpha, keep) of the sample
// supplied by the caller (as alpha_keep<0> computed them in the recompute).
template <int BLEND>
__device__ __forceinline__ void blend_adjoint_ak(float t_in, const f4& s, float step, float alpha, float keep,
                                                 f4& adj_acc, f4& adj_s, float& adj_step)
{
    const float dotc = adj_acc.x * s.x + adj_acc.y * s.y + adj_acc.z * s.z;
    const float adj_alpha = t_in * (adj_acc.w + dotc);
    const float k = alpha * t_in;
    adj_s.x = adj_acc.x * k;
    adj_s.y = adj_acc.y * k;
    adj_s.z = adj_acc.z * k;
    if constexpr (BLEND == 0) {
        adj_s.w = adj_alpha * step * keep;
        adj_step = adj_alpha * s.w * keep;
    } else {
        const bool lin = s.w * step < 1.0f;
        adj_step = lin ? adj_alpha * s.w : 0.0f;
        adj_s.w = lin ? adj_alpha * step : 0.0f;
    }
    adj_acc.w = adj_acc.w * keep - alpha * dotc;
}

template <int BLEND>
__device__ __forceinline__ void blend_adjoint_t(float t_in, const f4& s, float step, f4& adj_acc, f4& adj_s,
                                                float& adj_step)
{
    float alpha, keep;
    if constexpr (BLEND == 0) {
        alpha_keep<0>(s.w, step, alpha, keep);
    } else {
        alpha = tmin_(1.0f - 1e-5f, s.w * step);
        keep = 1.0f - alpha;
    }
    blend_adjoint_ak<BLEND>(t_in, s, step, alpha, keep, adj_acc, adj_s, adj_step);
}

// Blend adjoint for one sample given the accumulated alpha before it (a_in).
// adj_acc.xyz is constant along the ray; adj_acc.w is propagated in place.
// (renderer_blending.cuh:63-98 / :128-164 with a_in supplied instead of
//  reconstructed.)
template <int BLEND>
__device__ __forceinline__ void blend_adjoint(float a_in, const f4& s, float step, f4& adj_acc, f4& adj_s,
                                              float& adj_step)
{
    float alpha, e = 0.0f;
    if constexpr (BLEND == 0) {
        e = expf(-s.w * step);
        alpha = 1.0f - e;
    } else {
        alpha = tmin_(1.0f - 1e-5f, s.w * step);
    }
    const float tin = 1.0f - a_in;
    float adj_alpha = adj_acc.w * tin;
    float adj_ain = adj_acc.w * (1.0f - alpha);
    adj_alpha += adj_acc.x * (s.x - s.x * a_in) + adj_acc.y * (s.y - s.y * a_in) + adj_acc.z * (s.z - s.z * a_in);
    adj_ain += adj_acc.x * (-s.x * alpha) + adj_acc.y * (-s.y * alpha) + adj_acc.z * (-s.z * alpha);
    const float k = alpha * tin;
    adj_s.x = adj_acc.x * k;
    adj_s.y = adj_acc.y * k;
    adj_s.z = adj_acc.z * k;
    if constexpr (BLEND == 0) {
        adj_s.w = adj_alpha * step * e;
        adj_step = adj_alpha * s.w * e;
    } else {
        const bool lin = s.w * step < 1.0f;
        adj_step = lin ? adj_alpha * s.w : 0.0f;
        adj_s.w = lin ? adj_alpha * step : 0.0f;
    }
    adj_acc.w = adj_ain;
}

// Accumulated alpha before the sample from the one after it (reference
// inversion, renderer_blending.cuh:75-77 / :140-142).
template <int BLEND>
__device__ __forceinline__ f4 blend_invert(const f4& acc_out, const f4& s, float step)
{
    float alpha;
    if constexpr (BLEND == 0) alpha = 1.0f - expf(-s.w * step);
    else alpha = tmin_(1.0f - 1e-5f, s.w * step);
    const float a_in = (alpha - acc_out.w) / (alpha - 1.0f);
    const float tr = 1.0f - a_in;
    return mk4(acc_out.x - tr * s.x * alpha, acc_out.y - tr * s.y * alpha, acc_out.z - tr * s.z * alpha, a_in);
}

// ---------------------------------------------------------------------------
// TF adjoint.  The parameter-gradient sink is a policy object with
//   add(row, col, value)                       -- one entry
//   add_pair<POS>(r0, w0, r1, w1, adj, p0, p1) -- rows r0/r1 get adj*w0 / adj*w1
//                                                 in columns 0..3 (+ p0/p1 in
//                                                 column 4 when POS)
// so the same code feeds global atomics, LDS atomics or private arrays.
// Returns d(density).  (renderer_tf.cuh:75-104, 173-220, 315-400, 562-614)
// ---------------------------------------------------------------------------
template <int TFM, bool WANT_TF, class Sink, bool FAST = false>
__device__ __forceinline__ float tf_adjoint(const Tf& T, float density, const f4& adj, Sink& sink,
                                            const TfPre* pre = nullptr)
{
    if constexpr (TFM == 0) {
        const float dc = fclamp(density, 0.0f, 1.0f);
        const float sa = T.p[0], sc = T.p[T.sc];
        const float adj_dc = (adj.x * sc + adj.y * sc + adj.z * sc) + adj.w * sa;
        if constexpr (WANT_TF) {
            sink.add(0, 0, adj.w * dc);
            sink.add(0, 1, adj.x * dc + adj.y * dc + adj.z * dc);
        }
        return (density > 0.0f && density < 1.0f) ? adj_dc : 0.0f;
    } else if constexpr (TFM == 1) {
        const float d = density * float(T.R) - 0.5f;
        const int di = int(floorf(d));
        const float df = d - float(di);
        const int i0 = iclamp(di, 0, T.R - 1), i1 = iclamp(di + 1, 0, T.R - 1);
        const float* a = T.p + i0 * T.sr;
        const float* c = T.p + i1 * T.sr;
        const float adj_df = adj.x * (c[0] - a[0]) + adj.y * (c[T.sc] - a[T.sc]) + adj.z * (c[2 * T.sc] - a[2 * T.sc]) +
                             adj.w * (c[3 * T.sc] - a[3 * T.sc]);
        if constexpr (WANT_TF) sink.template add_pair<false>(i0, 1.0f - df, i1, df, adj, 0.0f, 0.0f);
        return adj_df * float(T.R);
    } else if constexpr (TFM == 2) {
        const TfLoc L = pre ? pre->L : tf_linear_locate<FAST>(T, density);
        const int i = L.i;
        const float* a = T.p + i * T.sr;
        const float* c = a + T.sr;
        const float p0 = L.p0, p1 = L.p1;
        const float d2 = fclamp(density, p0, p1);
        const float f = FAST ? (d2 - p0) * L.inv : (d2 - p0) / (p1 - p0);
        const float adj_f = adj.x * (c[0] - a[0]) + adj.y * (c[T.sc] - a[T.sc]) + adj.z * (c[2 * T.sc] - a[2 * T.sc]) +
                            adj.w * (c[3 * T.sc] - a[3 * T.sc]);
        if constexpr (WANT_TF) {
            float adj_p0, adj_p1;
            if constexpr (FAST) {
                const float q = adj_f * L.inv * L.inv;
                adj_p0 = q * (d2 - p1);
                adj_p1 = q * (p0 - d2);
            } else {
                const float den = (p1 - p0) * (p1 - p0);
                adj_p0 = adj_f * (d2 - p1) / den;
                adj_p1 = adj_f * (p0 - d2) / den;
            }
            sink.template add_pair<true>(i, 1.0f - f, i + 1, f, adj, adj_p0, adj_p1);
        }
        return FAST ? adj_f * L.inv : adj_f / (p1 - p0);  // passes through the clamp (:350-351)
    } else {
        float adj_d = 0.0f;
        auto comp = [&](int i, float n) __attribute__((always_inline)) {
            const float* r = T.p + i * T.sr;
            const float c0 = r[0], c1 = r[T.sc], c2 = r[2 * T.sc], c3 = r[3 * T.sc];
            const float mu = r[4 * T.sc], s = r[5 * T.sc];
            const float s2 = s * s;
            const float f1 = FAST ? (mu - density) * __builtin_amdgcn_rcpf(s2) * n : (mu - density) / s2 * n;
            const float g = adj.x * (c0 * f1) + adj.y * (c1 * f1) + adj.z * (c2 * f1) + adj.w * (c3 * f1);
            adj_d += g;
            if constexpr (WANT_TF) {
                const float f2 = FAST ? f1 * (mu - density) * __builtin_amdgcn_rcpf(s) : f1 * (mu - density) / s;
                sink.add(i, 0, adj.x * n);
                sink.add(i, 1, adj.y * n);
                sink.add(i, 2, adj.z * n);
                sink.add(i, 3, adj.w * n);
                sink.add(i, 4, -g);
                sink.add(i, 5, adj.x * (c0 * f2) + adj.y * (c1 * f2) + adj.z * (c2 * f2) + adj.w * (c3 * f2));
            }
        };
        int i0 = 0;
        if (pre) {
#pragma unroll
            for (int i = 0; i < kGaussCache; ++i)
                if (i < T.R) comp(i, pre->gn[i]);
            i0 = kGaussCache;
        }
        for (int i = i0; i < T.R; ++i) {
            const float* r = T.p + i * T.sr;
            comp(i, gauss_n<FAST>(density, r[4 * T.sc], r[5 * T.sc]));
        }
        return adj_d;
    }
}

struct SinkNone {
    __device__ __forceinline__ void add(int, int, float) {}
    template <bool POS> __device__ __forceinline__ void add_pair(int, float, int, float, const f4&, float, float) {}
};

// Nearest-neighbour voxel indices (renderer_interpolation.cuh:38-50)
__device__ __forceinline__ void nearest_index(f3 p, int X, int Y, int Z, int& ix, int& iy, int& iz)
{
    const float eps = 1e-7f;
    ix = int(roundf(fclamp(p.x, eps, float(X) - (1.0f + eps))));
    iy = int(roundf(fclamp(p.y, eps, float(Y) - (1.0f + eps))));
    iz = int(roundf(fclamp(p.z, eps, float(Z) - (1.0f + eps))));
}

// Trilinear adjoint of one sample: corner weights w[8] (times adj_d) and the
// position adjoint adj_cw - adj_fw (renderer_interpolation.cuh:360-387).
// `d` are the corner densities (needed only for the position adjoint).
template <bool WANT_POS>
__device__ __forceinline__ void tri_adjoint(const float d[8], float fx, float fy, float fz, float adj_d, float w[8],
                                            f3& adj_p)
{
    const float cx = fx, cy = fy, cz = fz;
    const float gx = 1.0f - fx, gy = 1.0f - fy, gz = 1.0f - fz;
    w[0] = adj_d * (gx * gy * gz);
    w[1] = adj_d * (gx * gy * cz);
    w[2] = adj_d * (gx * cy * gz);
    w[3] = adj_d * (gx * cy * cz);
    w[4] = adj_d * (cx * gy * gz);
    w[5] = adj_d * (cx * gy * cz);
    w[6] = adj_d * (cx * cy * gz);
    w[7] = adj_d * (cx * cy * cz);
    if constexpr (WANT_POS) {
        // d/dx of the trilinear form = sum over corners of +-(y-weight)(z-weight) d_k
        const float yz00 = gy * gz, yz01 = gy * cz, yz10 = cy * gz, yz11 = cy * cz;
        const float xz00 = gx * gz, xz01 = gx * cz, xz10 = cx * gz, xz11 = cx * cz;
        const float xy00 = gx * gy, xy01 = gx * cy, xy10 = cx * gy, xy11 = cx * cy;
        const float ddx = (d[4] - d[0]) * yz00 + (d[5] - d[1]) * yz01 + (d[6] - d[2]) * yz10 + (d[7] - d[3]) * yz11;
        const float ddy = (d[2] - d[0]) * xz00 + (d[3] - d[1]) * xz01 + (d[6] - d[4]) * xz10 + (d[7] - d[5]) * xz11;
        const float ddz = (d[1] - d[0]) * xy00 + (d[3] - d[2]) * xy01 + (d[5] - d[4]) * xy10 + (d[7] - d[6]) * xy11;
        adj_p = mk3(adj_d * ddx, adj_d * ddy, adj_d * ddz);
    }
}

}  // namespace dvr
