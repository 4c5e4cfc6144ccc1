// Pre-shaded volumes (renderer/renderer_interpolation.cuh:390-579,
// renderer/renderer_tf.cuh:818-854): the volume carries red, green, blue and
// absorption in its four batch slices; a sample's colour is the trilinear
// interpolation of each slice and the TF passes it on unchanged.  The render
// batch is a single image set (B taken from the camera, TF and step tensors).
//
// Forward and adjoint follow the reference's launch structure and adjoint
// method: one thread per pixel, grid-stride over pixels, accumulator
// reconstructed by inverting the blend step.  The position adjoint of a
// channel is the gradient of its trilinear form; the reference writes the z
// term into the y component (renderer_interpolation.cuh:557-562), which this
// kernel does not reproduce.
#include "dvr_device.hpp"
#include "dvr_launch.hpp"

namespace dvr {

constexpr int kInvalidCell4 = -0x40000000;

// the 8 corners of a cell in each of the 4 channels
__device__ __forceinline__ void cell4_load(const RenderParams& P, int ix, int iy, int iz, float d[4][8])
{
#pragma unroll
    for (int ch = 0; ch < 4; ++ch) cell_load(make_vol(P, ch), ix, iy, iz, d[ch]);
}

__device__ __forceinline__ f4 cell4_value(const float d[4][8], float fx, float fy, float fz)
{
    return mk4(tri_lerp(d[0], fx, fy, fz), tri_lerp(d[1], fx, fy, fz), tri_lerp(d[2], fx, fy, fz),
               tri_lerp(d[3], fx, fy, fz));
}

template <int BLEND>
__global__ __launch_bounds__(256) void k_forward_preshaded(RenderParams P, int total)
{
    const int stride = blockDim.x * gridDim.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const int b = i / (P.W * P.H);
        const int y = (i - b * P.W * P.H) / P.W;
        const int x = i - P.W * (y + P.H * b);
        const Ray r = setup_ray(P, x, y, b);
        float d[4][8];
        int cx = kInvalidCell4, cy = 0, cz = 0;
        f4 acc = mk4(0.f, 0.f, 0.f, 0.f);
        int n;
        for (n = 0; n < kMaxSamples; ++n) {
            const float t = r.t0 + float(n) * r.step;
            if (t > r.t1) break;
            if (acc.w > P.earlyOut) break;
            const f3 p = sample_pos(r, t);
            const int ix = int(p.x), iy = int(p.y), iz = int(p.z);
            if (ix != cx || iy != cy || iz != cz) {
                cell4_load(P, ix, iy, iz, d);
                cx = ix; cy = iy; cz = iz;
            }
            blend<BLEND>(acc, cell4_value(d, p.x - float(ix), p.y - float(iy), p.z - float(iz)), r.step);
        }
        float* o = P.color + b * P.colSb + y * P.colSy + x * P.colSx;
        o[0] = acc.x; o[P.colSc] = acc.y; o[2 * P.colSc] = acc.z; o[3 * P.colSc] = acc.w;
        P.term[b * P.termSb + y * P.termSy + x * P.termSx] = n;
    }
}

// corner gradients of the current cell, 4 channels, added to adj_volume with atomics
__device__ __forceinline__ void flush_cell4(const RenderParams& P, const AdjointParams& A, int cx, int cy, int cz,
                                            float w[4][8])
{
    const int lx = iclamp(cx, 0, P.X - 1), hx = iclamp(cx + 1, 0, P.X - 1);
    const int ly = iclamp(cy, 0, P.Y - 1), hy = iclamp(cy + 1, 0, P.Y - 1);
    const int lz = iclamp(cz, 0, P.Z - 1), hz = iclamp(cz + 1, 0, P.Z - 1);
    const int xs[2] = {lx * A.avSx, hx * A.avSx}, ys[2] = {ly * A.avSy, hy * A.avSy}, zs[2] = {lz * A.avSz, hz * A.avSz};
#pragma unroll
    for (int ch = 0; ch < 4; ++ch) {
        float* base = A.avol + ch * A.avSb;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (w[ch][k] != 0.0f) atomicAdd(base + xs[(k >> 2) & 1] + ys[(k >> 1) & 1] + zs[k & 1], w[ch][k]);
            w[ch][k] = 0.0f;
        }
    }
}

template <int BLEND>
__global__ __launch_bounds__(256) void k_adjoint_preshaded(RenderParams P, AdjointParams A, int total)
{
    const int stride = blockDim.x * gridDim.x;
    const bool wantVol = A.avol != nullptr;
    const bool wantPos = (A.ars != nullptr) || (A.ast != nullptr);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const int b = i / (P.W * P.H);
        const int y = (i - b * P.W * P.H) / P.W;
        const int x = i - P.W * (y + P.H * b);
        const Ray r = setup_ray(P, x, y, b);
        const float* fc = A.fcol + b * A.fcSb + y * A.fcSy + x * A.fcSx;
        f4 acc = mk4(fc[0], fc[A.fcSc], fc[2 * A.fcSc], fc[3 * A.fcSc]);
        const int n = min(A.fterm[b * A.ftSb + y * A.ftSy + x * A.ftSx], kMaxSamples);
        const float* ac = A.acol + b * A.acSb + y * A.acSy + x * A.acSx;
        f4 adj_acc = mk4(ac[0], ac[A.acSc], ac[2 * A.acSc], ac[3 * A.acSc]);
        float adj_step = 0.0f;
        f3 adj_start = mk3(0.f, 0.f, 0.f), adj_dir = mk3(0.f, 0.f, 0.f);
        float d[4][8], w[4][8];
#pragma unroll
        for (int ch = 0; ch < 4; ++ch)
#pragma unroll
            for (int k = 0; k < 8; ++k) w[ch][k] = 0.0f;
        int cx = kInvalidCell4, cy = 0, cz = 0;
        for (int s_i = n - 1; s_i >= 0; --s_i) {
            const float t = r.t0 + float(s_i) * r.step;
            const f3 p = sample_pos(r, t);
            const int ix = int(p.x), iy = int(p.y), iz = int(p.z);
            if (ix != cx || iy != cy || iz != cz) {
                if (wantVol && cx != kInvalidCell4) flush_cell4(P, A, cx, cy, cz, w);
                cell4_load(P, ix, iy, iz, d);
                cx = ix; cy = iy; cz = iz;
            }
            const float fx = p.x - float(ix), fy = p.y - float(iy), fz = p.z - float(iz);
            const f4 s = cell4_value(d, fx, fy, fz);
            const f4 accIn = blend_invert<BLEND>(acc, s, r.step);
            f4 adj_s;
            float adj_st;
            blend_adjoint<BLEND>(accIn.w, s, r.step, adj_acc, adj_s, adj_st);
            acc = accIn;
            // the TF is the identity on the colour: adj_colour of the fetch = adj_s
            const float as[4] = {adj_s.x, adj_s.y, adj_s.z, adj_s.w};
            f3 adj_p = mk3(0.f, 0.f, 0.f);
#pragma unroll
            for (int ch = 0; ch < 4; ++ch) {
                float wc[8];
                f3 ap;
                if (wantPos) tri_adjoint<true>(d[ch], fx, fy, fz, as[ch], wc, ap);
                else tri_adjoint<false>(d[ch], fx, fy, fz, as[ch], wc, ap);
                if (wantPos) adj_p = add3(adj_p, ap);
#pragma unroll
                for (int k = 0; k < 8; ++k) w[ch][k] += wc[k];
            }
            if (wantPos) {
                const f3 adj_w = mul3(adj_p, r.invVoxel);
                adj_start = add3(adj_start, adj_w);
                adj_dir = add3(adj_dir, scl3(adj_w, t));
                adj_step += dot3(adj_w, r.dir) * float(s_i);
            }
            adj_step += adj_st;
        }
        if (wantVol && cx != kInvalidCell4) flush_cell4(P, A, cx, cy, cz, w);
        if (A.ast) {
            float* o = A.ast + b * A.astSb + y * A.astSy + x * A.astSx;
            if (A.stepBroadcast) atomicAdd(o, adj_step);
            else *o = adj_step;
        }
        if (A.ars) {
            float* os = A.ars + b * A.arsSb + y * A.arsSy + x * A.arsSx;
            float* od = A.ard + b * A.ardSb + y * A.ardSy + x * A.ardSx;
            if (A.camBroadcast) {
                atomicAdd(os, adj_start.x); atomicAdd(os + A.arsSc, adj_start.y); atomicAdd(os + 2 * A.arsSc, adj_start.z);
                atomicAdd(od, adj_dir.x); atomicAdd(od + A.ardSc, adj_dir.y); atomicAdd(od + 2 * A.ardSc, adj_dir.z);
            } else {
                os[0] = adj_start.x; os[A.arsSc] = adj_start.y; os[2 * A.arsSc] = adj_start.z;
                od[0] = adj_dir.x; od[A.ardSc] = adj_dir.y; od[2 * A.ardSc] = adj_dir.z;
            }
        }
    }
}

// The forward march gathers 32 floats per new cell: at most two 256-thread
// workgroups per CU keep the rays in flight within what L1 holds (C2 scene
// pre-shaded, 4 x 128^3: 3.04 -> 2.72 ms).  The adjoint is bound by its
// volume-gradient atomics and keeps the register-limited occupancy.
constexpr int kPreshadedFwdWgPerCu = 2;

static int preshaded_grid(int total)
{
    const int want = (total + 255) / 256;
    return want < 4096 ? want : 4096;
}

hipError_t launch_forward_preshaded(const RenderParams& P, int blend, hipStream_t s)
{
    const int total = P.W * P.H * P.B;
    if (total == 0) return hipSuccess;
    const size_t lds = occupancy_lds(0, kPreshadedFwdWgPerCu);
    if (blend == 0) hipLaunchKernelGGL(k_forward_preshaded<0>, dim3(preshaded_grid(total)), dim3(256), lds, s, P, total);
    else hipLaunchKernelGGL(k_forward_preshaded<1>, dim3(preshaded_grid(total)), dim3(256), lds, s, P, total);
    return hipGetLastError();
}

hipError_t launch_adjoint_preshaded(const RenderParams& P, const AdjointParams& A, int blend, hipStream_t s)
{
    const int total = P.W * P.H * P.B;
    if (total == 0) return hipSuccess;
    if (blend == 0) hipLaunchKernelGGL(k_adjoint_preshaded<0>, dim3(preshaded_grid(total)), dim3(256), 0, s, P, A, total);
    else hipLaunchKernelGGL(k_adjoint_preshaded<1>, dim3(preshaded_grid(total)), dim3(256), 0, s, P, A, total);
    return hipGetLastError();
}

}  // namespace dvr
