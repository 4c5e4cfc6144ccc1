// Camera ray generation and its adjoint (renderer/camera_impl.cu:17-93,
// :134-174, :302-357).  The adjoint's per-camera segmented sum (cub in the
// reference) is fused into the per-pixel kernel: wave reduction, LDS
// reduction across the workgroup, one atomic per value and workgroup.
#include "dvr_device.hpp"
#include "dvr_launch.hpp"

namespace dvr {

struct Viewport {
    f3 eye, right, up;
};

__device__ __forceinline__ Viewport load_viewport(const float* v, int sr, int sc)
{
    Viewport m;
    m.eye = mk3(v[0], v[sc], v[2 * sc]);
    m.right = mk3(v[sr], v[sr + sc], v[sr + 2 * sc]);
    m.up = mk3(v[2 * sr], v[2 * sr + sc], v[2 * sr + 2 * sc]);
    return m;
}

__global__ __launch_bounds__(256) void k_generate_rays(const float* vp, int vSb, int vSr, int vSc, int H, int W,
                                                       float tanX, float tanY, float* rs, int4 rsS, float* rd,
                                                       int4 rdS)
{
    const int b = blockIdx.y;
    const int pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= H * W) return;
    const int y = pix / W, x = pix - y * W;
    const Viewport m = load_viewport(vp + b * vSb, vSr, vSc);
    const f3 front = cross3(m.up, m.right);
    const float fx = 2.0f * (x + 0.5f) / float(W) - 1.0f;
    const float fy = 2.0f * (y + 0.5f) / float(H) - 1.0f;
    const f3 d = normalize3(add3(add3(front, scl3(m.right, fx * tanX)), scl3(m.up, fy * tanY)));
    float* s = rs + b * rsS.x + y * rsS.y + x * rsS.z;
    float* o = rd + b * rdS.x + y * rdS.y + x * rdS.z;
    s[0] = m.eye.x; s[rsS.w] = m.eye.y; s[2 * rsS.w] = m.eye.z;
    o[0] = d.x; o[rdS.w] = d.y; o[2 * rdS.w] = d.z;
}

__global__ void k_zero_viewport(float* out, int oSb, int oSr, int oSc, int B)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B * 9) return;
    const int b = i / 9, k = i - 9 * b;
    out[b * oSb + (k / 3) * oSr + (k % 3) * oSc] = 0.0f;
}

__global__ __launch_bounds__(256) void k_generate_rays_adjoint(const float* vp, int vSb, int vSr, int vSc, int H,
                                                               int W, float tanX, float tanY, const float* ars,
                                                               int4 arsS, const float* ard, int4 ardS, float* out,
                                                               int oSb, int oSr, int oSc)
{
    __shared__ float red[4][9];
    const int b = blockIdx.y;
    const int pix = blockIdx.x * blockDim.x + threadIdx.x;
    float v[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) v[k] = 0.0f;
    if (pix < H * W) {
        const int y = pix / W, x = pix - y * W;
        const Viewport m = load_viewport(vp + b * vSb, vSr, vSc);
        const f3 front = cross3(m.up, m.right);
        const float fx = 2.0f * (x + 0.5f) / float(W) - 1.0f;
        const float fy = 2.0f * (y + 0.5f) / float(H) - 1.0f;
        const f3 d = add3(add3(front, scl3(m.right, fx * tanX)), scl3(m.up, fy * tanY));
        const float* s = ars + b * arsS.x + y * arsS.y + x * arsS.z;
        const float* o = ard + b * ardS.x + y * ardS.y + x * ardS.z;
        const f3 ae = mk3(s[0], s[arsS.w], s[2 * arsS.w]);
        const f3 adn = mk3(o[0], o[ardS.w], o[2 * ardS.w]);
        const f3 ad = adj_normalize(d, adn);
        f3 ar = scl3(ad, fx * tanX);
        f3 au = scl3(ad, fy * tanY);
        adj_cross(m.up, m.right, ad, au, ar);
        v[0] = ae.x; v[1] = ae.y; v[2] = ae.z;
        v[3] = ar.x; v[4] = ar.y; v[5] = ar.z;
        v[6] = au.x; v[7] = au.y; v[8] = au.z;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const float s = wave_sum(v[k]);
        if (lane == 0) red[wave][k] = s;
    }
    __syncthreads();
    if (threadIdx.x < 9) {
        const int k = threadIdx.x;
        const float s = red[0][k] + red[1][k] + red[2][k] + red[3][k];
        atomicAdd(out + b * oSb + (k / 3) * oSr + (k % 3) * oSc, s);
    }
}

hipError_t launch_generate_rays(const float* viewport, int vSb, int vSr, int vSc, int B, int H, int W, float fovY,
                                float* rs, const int rsS[4], float* rd, const int rdS[4], hipStream_t s)
{
    if (B == 0 || H == 0 || W == 0) return hipSuccess;
    // tangents evaluated in float on the host, as camera_impl.cu:74-76
    const float fovX = fovY * float(W) / float(H);
    const float tanX = tanf(fovX / 2), tanY = tanf(fovY / 2);
    dim3 grid((H * W + 255) / 256, B);
    hipLaunchKernelGGL(k_generate_rays, grid, dim3(256), 0, s, viewport, vSb, vSr, vSc, H, W, tanX, tanY, rs,
                       make_int4(rsS[0], rsS[1], rsS[2], rsS[3]), rd, make_int4(rdS[0], rdS[1], rdS[2], rdS[3]));
    return hipGetLastError();
}

hipError_t launch_generate_rays_adjoint(const float* viewport, int vSb, int vSr, int vSc, int B, int H, int W,
                                        float fovY, const float* ars, const int arsS[4], const float* ard,
                                        const int ardS[4], float* out, int oSb, int oSr, int oSc, hipStream_t s)
{
    if (B == 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero_viewport, dim3((B * 9 + 255) / 256), dim3(256), 0, s, out, oSb, oSr, oSc, B);
    if (H == 0 || W == 0) return hipGetLastError();
    const float fovX = fovY * float(W) / float(H);
    const float tanX = tanf(fovX / 2), tanY = tanf(fovY / 2);
    dim3 grid((H * W + 255) / 256, B);
    hipLaunchKernelGGL(k_generate_rays_adjoint, grid, dim3(256), 0, s, viewport, vSb, vSr, vSc, H, W, tanX, tanY, ars,
                       make_int4(arsS[0], arsS[1], arsS[2], arsS[3]), ard, make_int4(ardS[0], ardS[1], ardS[2], ardS[3]),
                       out, oSb, oSr, oSc);
    return hipGetLastError();
}

}  // namespace dvr
