// Kernels around the batched launch: clearing gradient buffers and counters
// (a kernel, so that HIP graphs hold no memset nodes) and a multi-block Adam
// step (torch.optim.Adam semantics, torch/optim/adam.py single-tensor path:
// exp_avg = lerp(exp_avg, g, 1-b1); exp_avg_sq = b2 exp_avg_sq + (1-b2) g^2;
// p -= lr/(1-b1^t) * exp_avg / (sqrt(exp_avg_sq)/sqrt(1-b2^t) + eps)).
// The step count t lives in device memory.
#include "dvr_device.hpp"
#include "dvr_launch.hpp"

namespace dvr {

// blockIdx.y = segment; grid-stride over its elements
__global__ __launch_bounds__(256) void k_zero_segments(const ZeroSeg* segs)
{
    const ZeroSeg s = segs[blockIdx.y];
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < s.n; i += (long long)gridDim.x * 256)
        s.p[i] = 0.0f;
}

hipError_t launch_zero_segments(const ZeroSeg* dSegs, int nSegs, long long maxN, hipStream_t s)
{
    if (nSegs <= 0) return hipSuccess;
    long long bx = (maxN + 255) / 256;
    if (bx > 1024) bx = 1024;
    if (bx < 1) bx = 1;
    hipLaunchKernelGGL(k_zero_segments, dim3(unsigned(bx), unsigned(nSegs)), dim3(256), 0, s, dSegs);
    return hipGetLastError();
}

// Advances the step counter of an Adam group once per step (one thread), so
// that a captured sequence of steps keeps counting across graph replays.
__global__ void k_adam_tick(long long* step) { *step += 1; }

// Bias corrections are formed in double from the step counter and rounded to
// float, as torch rounds its Python-float step size and sqrt(bias_correction2)
// when it hands them to the float kernels.
__global__ __launch_bounds__(256) void k_adam(const AdamSeg* segs, const long long* step, double lr, double b1,
                                              double b2, double eps)
{
    __shared__ float s_c[2];
    if (threadIdx.x == 0) {
        const double t = double(*step);
        const double bc1 = 1.0 - pow(b1, t), bc2 = 1.0 - pow(b2, t);
        s_c[0] = float(lr / bc1);
        s_c[1] = float(sqrt(bc2));
    }
    __syncthreads();
    const float stepSize = s_c[0], bias2Sqrt = s_c[1];
    const float w1 = float(1.0 - b1), fb2 = float(b2), w2 = float(1.0 - b2), feps = float(eps);
    const AdamSeg a = segs[blockIdx.y];
    const long long n = a.rows * a.cols;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        long long o = i;
        if (a.stride != a.cols) {
            const long long r = i / a.cols;
            o = r * a.stride + (i - r * a.cols);
        }
        const float g = a.grad[o];
        const float m0 = a.m[o];
        const float m = m0 + w1 * (g - m0);  // lerp(exp_avg, g, 1 - b1)
        const float v = a.v[o] * fb2 + (w2 * g) * g;
        a.m[o] = m;
        a.v[o] = v;
        const float denom = sqrtf(v) / bias2Sqrt + feps;
        float p = a.p[o] + (-stepSize) * (m / denom);
        if (a.lo <= a.hi) p = fminf(fmaxf(p, a.lo), a.hi);
        a.p[o] = p;
    }
}

hipError_t launch_adam(const AdamSeg* dSegs, int nSegs, long long maxN, long long* dStep, double lr, double b1,
                       double b2, double eps, hipStream_t s)
{
    hipLaunchKernelGGL(k_adam_tick, dim3(1), dim3(1), 0, s, dStep);
    if (nSegs <= 0) return hipGetLastError();
    long long bx = (maxN + 255) / 256;
    if (bx > 2048) bx = 2048;
    if (bx < 1) bx = 1;
    hipLaunchKernelGGL(k_adam, dim3(unsigned(bx), unsigned(nSegs)), dim3(256), 0, s, dSegs, dStep, lr, b1, b2, eps);
    return hipGetLastError();
}

}  // namespace dvr
