// "Plain" kernels: the reference's launch structure on MI355X.
// One thread per pixel, pixels in row order (x fastest), 1-D grid-stride
// loop (renderer/renderer_utils.cuh:303-310, renderer/renderer_cuda.cpp:85-118),
// adjoint by inverting the blend step (renderer/renderer_kernels.cuh:315-444).
// These are the per-view baseline the batched checkpointed path is measured
// against, and the path taken for inputs the tiled kernels do not cover.
#include "dvr_device.hpp"
#include "dvr_launch.hpp"

namespace dvr {

// ---------------------------------------------------------------------------
// forward
// ---------------------------------------------------------------------------
template <int FILTER, int TFM, int BLEND>
__global__ __launch_bounds__(256) void k_forward_plain(RenderParams P, int total)
{
    const int stride = blockDim.x * gridDim.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const int b = i / (P.W * P.H);
        const int y = (i - b * P.W * P.H) / P.W;
        const int x = i - P.W * (y + P.H * b);
        const Ray r = setup_ray(P, x, y, b);
        const Vol V = make_vol(P, b);
        const Tf T{P.tf + b * P.tfSb, P.R, P.C, P.tfSr, P.tfSc};
        Cell c;
        cell_invalidate(c);
        f4 acc = mk4(0.f, 0.f, 0.f, 0.f);
        int n;
        for (n = 0; n < kMaxSamples; ++n) {
            const float t = r.t0 + float(n) * r.step;
            if (t > r.t1) break;
            if (acc.w > P.earlyOut) break;
            const float d = fetch<FILTER>(V, sample_pos(r, t), c);
            blend<BLEND>(acc, tf_eval<TFM>(T, d), r.step);
        }
        float* o = P.color + b * P.colSb + y * P.colSy + x * P.colSx;
        o[0] = acc.x; o[P.colSc] = acc.y; o[2 * P.colSc] = acc.z; o[3 * P.colSc] = acc.w;
        P.term[b * P.termSb + y * P.termSy + x * P.termSx] = n;
    }
}

// ---------------------------------------------------------------------------
// TF-gradient sinks of the plain adjoint
// ---------------------------------------------------------------------------
// immediate: one global atomic per entry and sample (renderer_tf.cuh:208-216)
struct SinkGlobalAtomic {
    float* p; int sr, sc;
    __device__ void add(int r, int c, float v) { atomicAdd(p + r * sr + c * sc, v); }
    template <bool POS>
    __device__ void add_pair(int r0, float w0, int r1, float w1, const f4& a, float p0, float p1)
    {
        add(r0, 0, a.x * w0); add(r0, 1, a.y * w0); add(r0, 2, a.z * w0); add(r0, 3, a.w * w0);
        if constexpr (POS) add(r0, 4, p0);
        add(r1, 0, a.x * w1); add(r1, 1, a.y * w1); add(r1, 2, a.z * w1); add(r1, 3, a.w * w1);
        if constexpr (POS) add(r1, 4, p1);
    }
};

// delayed: a private R*C slice per thread in LDS, entry-major
// (renderer_tf.cuh:196-206 with renderer_cuda.cpp:273-295)
struct SinkThreadPrivate {
    float* p; int C, stride;
    __device__ void add(int r, int c, float v) { p[(r * C + c) * stride] += v; }
    template <bool POS>
    __device__ void add_pair(int r0, float w0, int r1, float w1, const f4& a, float p0, float p1)
    {
        add(r0, 0, a.x * w0); add(r0, 1, a.y * w0); add(r0, 2, a.z * w0); add(r0, 3, a.w * w0);
        if constexpr (POS) add(r0, 4, p0);
        add(r1, 0, a.x * w1); add(r1, 1, a.y * w1); add(r1, 2, a.z * w1); add(r1, 3, a.w * w1);
        if constexpr (POS) add(r1, 4, p1);
    }
};

// ---------------------------------------------------------------------------
// adjoint by inversion.  TFACC: 0 = no TF gradient, 1 = immediate, 2 = delayed
// ---------------------------------------------------------------------------
template <int FILTER, int TFM, int BLEND, int TFACC>
__global__ __launch_bounds__(256) void k_adjoint_plain(RenderParams P, AdjointParams A, int total)
{
    extern __shared__ float s_priv[];
    const int stride = blockDim.x * gridDim.x;
    const int lane = threadIdx.x & 63;
    const bool wantVol = A.avol != nullptr;
    const bool wantPos = (A.ars != nullptr) || (A.ast != nullptr);
    // wave-uniform trip count so the delayed reduction sees whole waves
    for (int base = blockIdx.x * blockDim.x; base < total; base += stride) {
        const int i = base + threadIdx.x;
        const bool active = i < total;
        const int b = active ? i / (P.W * P.H) : 0;
        if constexpr (TFACC == 2) {
            for (int e = 0; e < P.R * P.C; ++e) s_priv[e * blockDim.x + threadIdx.x] = 0.0f;
        }
        if (active) {
            const int y = (i - b * P.W * P.H) / P.W;
            const int x = i - P.W * (y + P.H * b);
            const Ray r = setup_ray(P, x, y, b);
            const Vol V = make_vol(P, b);
            const Tf T{P.tf + b * P.tfSb, P.R, P.C, P.tfSr, P.tfSc};
            const float* fc = A.fcol + b * A.fcSb + y * A.fcSy + x * A.fcSx;
            f4 acc = mk4(fc[0], fc[A.fcSc], fc[2 * A.fcSc], fc[3 * A.fcSc]);
            const int n = min(A.fterm[b * A.ftSb + y * A.ftSy + x * A.ftSx], kMaxSamples);
            const float* ac = A.acol + b * A.acSb + y * A.acSy + x * A.acSx;
            f4 adj_acc = mk4(ac[0], ac[A.acSc], ac[2 * A.acSc], ac[3 * A.acSc]);
            float* avol = wantVol ? A.avol + b * A.avSb : nullptr;

            float adj_step = 0.0f;
            f3 adj_start = mk3(0.f, 0.f, 0.f), adj_dir = mk3(0.f, 0.f, 0.f);
            // trilinear: cell cache of densities + adjoint weights, flushed on
            // cell change (renderer_interpolation.cuh:299-342)
            Cell c;
            cell_invalidate(c);
            float wacc[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) wacc[k] = 0.0f;
            int nearestLast = -1;
            float nearestAcc = 0.0f;

            SinkGlobalAtomic sg{A.atf ? A.atf + b * A.atfSb : nullptr, A.atfSr, A.atfSc};
            SinkThreadPrivate sp{s_priv + threadIdx.x, P.C, (int)blockDim.x};

            for (int s_i = n - 1; s_i >= 0; --s_i) {
                const float t = r.t0 + float(s_i) * r.step;
                const f3 p = sample_pos(r, t);
                float d;
                float fx = 0.f, fy = 0.f, fz = 0.f;
                int noff = 0;
                if constexpr (FILTER == 0) {
                    noff = nearest_offset(V, p);
                    d = gld(V.p, unsigned(noff));
                } else {
                    const int ix = int(p.x), iy = int(p.y), iz = int(p.z);
                    if (ix != c.ix || iy != c.iy || iz != c.iz) {
                        if (wantVol && c.ix != -0x40000000) {
                            const int lx = iclamp(c.ix, 0, P.X - 1), hx = iclamp(c.ix + 1, 0, P.X - 1);
                            const int ly = iclamp(c.iy, 0, P.Y - 1), hy = iclamp(c.iy + 1, 0, P.Y - 1);
                            const int lz = iclamp(c.iz, 0, P.Z - 1), hz = iclamp(c.iz + 1, 0, P.Z - 1);
                            const int xs[2] = {lx * A.avSx, hx * A.avSx}, ys[2] = {ly * A.avSy, hy * A.avSy},
                                      zs[2] = {lz * A.avSz, hz * A.avSz};
#pragma unroll
                            for (int k = 0; k < 8; ++k) {
                                atomicAdd(avol + xs[(k >> 2) & 1] + ys[(k >> 1) & 1] + zs[k & 1], wacc[k]);
                                wacc[k] = 0.0f;
                            }
                        }
                        cell_load(V, ix, iy, iz, c.d);
                        c.ix = ix; c.iy = iy; c.iz = iz;
                    }
                    fx = p.x - float(ix); fy = p.y - float(iy); fz = p.z - float(iz);
                    d = tri_lerp(c.d, fx, fy, fz);
                }
                const f4 s = tf_eval<TFM>(T, d);
                const f4 accIn = blend_invert<BLEND>(acc, s, r.step);
                f4 adj_s;
                float adj_st;
                blend_adjoint<BLEND>(accIn.w, s, r.step, adj_acc, adj_s, adj_st);
                acc = accIn;
                float adj_d;
                if constexpr (TFACC == 1) adj_d = tf_adjoint<TFM, true>(T, d, adj_s, sg);
                else if constexpr (TFACC == 2) adj_d = tf_adjoint<TFM, true>(T, d, adj_s, sp);
                else { SinkNone sn; adj_d = tf_adjoint<TFM, false>(T, d, adj_s, sn); }
                f3 adj_p = mk3(0.f, 0.f, 0.f);
                if constexpr (FILTER == 0) {
                    if (wantVol) {
                        int gx, gy, gz;
                        nearest_index(p, P.X, P.Y, P.Z, gx, gy, gz);
                        const int goff = gx * A.avSx + gy * A.avSy + gz * A.avSz;
                        if (goff != nearestLast) {
                            if (nearestLast >= 0) atomicAdd(avol + nearestLast, nearestAcc);
                            nearestLast = goff;
                            nearestAcc = 0.0f;
                        }
                        nearestAcc += adj_d;
                    }
                } else {
                    float w[8];
                    if (wantPos) tri_adjoint<true>(c.d, fx, fy, fz, adj_d, w, adj_p);
                    else tri_adjoint<false>(c.d, fx, fy, fz, adj_d, w, adj_p);
#pragma unroll
                    for (int k = 0; k < 8; ++k) wacc[k] += w[k];
                }
                if (wantPos) {
                    const f3 adj_w = mul3(adj_p, r.invVoxel);
                    adj_start = add3(adj_start, adj_w);
                    adj_dir = add3(adj_dir, scl3(adj_w, t));
                    adj_step += dot3(adj_w, r.dir) * float(s_i);
                }
                adj_step += adj_st;
            }
            // flush the last cell (destructor in the reference, :324)
            if constexpr (FILTER == 0) {
                if (wantVol && nearestLast >= 0) atomicAdd(avol + nearestLast, nearestAcc);
            } else {
                if (wantVol && c.ix != -0x40000000) {
                    const int lx = iclamp(c.ix, 0, P.X - 1), hx = iclamp(c.ix + 1, 0, P.X - 1);
                    const int ly = iclamp(c.iy, 0, P.Y - 1), hy = iclamp(c.iy + 1, 0, P.Y - 1);
                    const int lz = iclamp(c.iz, 0, P.Z - 1), hz = iclamp(c.iz + 1, 0, P.Z - 1);
                    const int xs[2] = {lx * A.avSx, hx * A.avSx}, ys[2] = {ly * A.avSy, hy * A.avSy},
                              zs[2] = {lz * A.avSz, hz * A.avSz};
#pragma unroll
                    for (int k = 0; k < 8; ++k) atomicAdd(avol + xs[(k >> 2) & 1] + ys[(k >> 1) & 1] + zs[k & 1], wacc[k]);
                }
            }
            // per-ray outputs (renderer_kernels.cuh:416-439)
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
        if constexpr (TFACC == 2) {
            // wave reduction of every private entry, leader adds (renderer_tf.cuh:221-245).
            // A wave may straddle two batch entries; reduce per batch index.
            const int b0 = __shfl(b, 0, 64);
            const bool uniformB = __all(b == b0 || !active);
            for (int e = 0; e < P.R * P.C; ++e) {
                const float v = s_priv[e * blockDim.x + threadIdx.x];
                const int rr = e / P.C, cc = e - rr * P.C;
                if (uniformB) {
                    const float sum = wave_sum(v);
                    if (lane == 0) atomicAdd(A.atf + b0 * A.atfSb + rr * A.atfSr + cc * A.atfSc, sum);
                } else if (active && v != 0.0f) {
                    atomicAdd(A.atf + b * A.atfSb + rr * A.atfSr + cc * A.atfSc, v);
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static int plain_grid(int total, int block)
{
    // like the reference: min(ceil(total/block), occupancy-limited grid)
    const int want = (total + block - 1) / block;
    const int cap = 256 * 16;  // 256 CUs x 16 resident blocks
    return want < cap ? want : cap;
}

template <int FILTER, int TFM, int BLEND>
static hipError_t fwd_plain_t(const RenderParams& P, hipStream_t s)
{
    const int total = P.W * P.H * P.B;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL((k_forward_plain<FILTER, TFM, BLEND>), dim3(plain_grid(total, 256)), dim3(256), 0, s, P, total);
    return hipGetLastError();
}

template <int FILTER, int TFM, int BLEND>
static hipError_t adj_plain_t(const RenderParams& P, const AdjointParams& A, hipStream_t s)
{
    const int total = P.W * P.H * P.B;
    if (total == 0) return hipSuccess;
    if (!A.atf) {
        hipLaunchKernelGGL((k_adjoint_plain<FILTER, TFM, BLEND, 0>), dim3(plain_grid(total, 256)), dim3(256), 0, s, P, A, total);
        return hipGetLastError();
    }
    if (A.tfDelayed) {
        // largest block whose private slices fit into 64 KiB of LDS
        const size_t per = sizeof(float) * size_t(P.R) * P.C;
        int block = 256;
        while (block > 64 && per * block > 65536) block >>= 1;
        if (per * block <= 65536) {
            hipLaunchKernelGGL((k_adjoint_plain<FILTER, TFM, BLEND, 2>), dim3(plain_grid(total, block)), dim3(block),
                               per * block, s, P, A, total);
            return hipGetLastError();
        }
        // too large for per-thread slices: immediate accumulation
    }
    hipLaunchKernelGGL((k_adjoint_plain<FILTER, TFM, BLEND, 1>), dim3(plain_grid(total, 256)), dim3(256), 0, s, P, A, total);
    return hipGetLastError();
}

hipError_t launch_forward_plain(const RenderParams& P, int filter, int tfm, int blend, hipStream_t s)
{
    DVR_DISPATCH3(filter, tfm, blend, return fwd_plain_t, (P, s));
    return hipErrorInvalidValue;
}

hipError_t launch_adjoint_plain(const RenderParams& P, const AdjointParams& A, int filter, int tfm, int blend,
                                hipStream_t s)
{
    DVR_DISPATCH3(filter, tfm, blend, return adj_plain_t, (P, A, s));
    return hipErrorInvalidValue;
}

}  // namespace dvr
