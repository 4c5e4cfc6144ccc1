// Dispatcher over the batched persistent kernel instance slices (kernels in
// dvr_batched_impl.hpp).
#include "dvr_launch.hpp"

namespace dvr {

#define DVR_DECL_INST(F, B, K)                                                                                  \
    hipError_t batched_inst_##F##_##B##_##K(const BatchParams&, int, int, int, size_t, hipStream_t, int*);
#define DVR_DECL_K(F, B) DVR_DECL_INST(F, B, 4) DVR_DECL_INST(F, B, 8) DVR_DECL_INST(F, B, 16) DVR_DECL_INST(F, B, 32)
DVR_DECL_K(0, 0)
DVR_DECL_K(0, 1)
DVR_DECL_K(1, 0)
DVR_DECL_K(1, 1)
#undef DVR_DECL_K
#undef DVR_DECL_INST

hipError_t launch_batched(const BatchParams& P, int filter, int tfm, int blend, int K, int gm, int grid, size_t lds,
                          hipStream_t s, int* occ)
{
    if (filter < 0 || filter > 1 || blend < 0 || blend > 1) return hipErrorInvalidValue;
    if (gm == 0) K = 16;  // the forward-only instances live in the K = 16 slices
#define DVR_CALL(F, B, KK) return batched_inst_##F##_##B##_##KK(P, tfm, gm, grid, lds, s, occ)
#define DVR_SEL_K(F, B)                      \
    switch (K) {                             \
    case 4: DVR_CALL(F, B, 4);               \
    case 8: DVR_CALL(F, B, 8);               \
    case 16: DVR_CALL(F, B, 16);             \
    case 32: DVR_CALL(F, B, 32);             \
    default: return hipErrorInvalidValue;    \
    }
    if (filter == 0 && blend == 0) { DVR_SEL_K(0, 0) }
    if (filter == 0 && blend == 1) { DVR_SEL_K(0, 1) }
    if (filter == 1 && blend == 0) { DVR_SEL_K(1, 0) }
    { DVR_SEL_K(1, 1) }
#undef DVR_SEL_K
#undef DVR_CALL
}

// Adds each volume's bricked gradient scratch (2 x 2 x 4 voxels per 64-byte
// chunk, bricks z-fastest) into the caller's contiguous X x Y x Z gradient: a
// thread per brick, consecutive threads on consecutive z bricks of one (x, y)
// brick column, so reads and the four 16-byte row segments it updates are
// contiguous across the wave.  blockIdx.y = volume.
__global__ __launch_bounds__(256) void k_debrick(const DebrickSeg* segs)
{
    const DebrickSeg S = segs[blockIdx.y];
    const int BY = (S.Y + 1) >> 1, BZ = (S.Z + 3) >> 2;
    const long long nb = (long long)((S.X + 1) >> 1) * BY * BZ;
    const bool vec = (S.Z & 3) == 0;
    for (long long b = (long long)blockIdx.x * 256 + threadIdx.x; b < nb; b += (long long)gridDim.x * 256) {
        const int zb = int(b % BZ);
        const long long r = b / BZ;
        const int yb = int(r % BY), xb = int(r / BY);
        const float4* src = reinterpret_cast<const float4*>(S.brick) + b * 4;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int x = 2 * xb + (q >> 1), y = 2 * yb + (q & 1);
            if (x >= S.X || y >= S.Y) continue;
            const float4 v = src[q];
            float* g = S.adj + ((size_t)x * S.Y + y) * S.Z + 4 * zb;
            if (vec) {
                float4 o = *reinterpret_cast<float4*>(g);
                o.x += v.x; o.y += v.y; o.z += v.z; o.w += v.w;
                *reinterpret_cast<float4*>(g) = o;
            } else {
                const float vv[4] = {v.x, v.y, v.z, v.w};
                for (int k = 0; k < 4 && 4 * zb + k < S.Z; ++k) g[k] += vv[k];
            }
        }
    }
}

hipError_t launch_debrick(const DebrickSeg* dSegs, int nSegs, long long maxBricks, hipStream_t s)
{
    if (nSegs <= 0) return hipSuccess;
    long long bx = (maxBricks + 255) / 256;
    if (bx > 2048) bx = 2048;
    if (bx < 1) bx = 1;
    hipLaunchKernelGGL(k_debrick, dim3(unsigned(bx), unsigned(nSegs)), dim3(256), 0, s, dSegs);
    return hipGetLastError();
}

}  // namespace dvr
