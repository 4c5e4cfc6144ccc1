// Batched persistent forward + loss + checkpointed adjoint over many views of
// many volumes (each view its own camera, image size, step, volume and TF).
//
// One launch covers the whole batch.  A grid sized to fill the chip once
// (occupancy x CUs) loops over workgroup tiles popped from per-XCD queues:
// tiles are numbered view after view and cut into eight contiguous ranges,
// workgroup w serves XCD w % 8 first (workgroups are dispatched round-robin
// over the XCDs), so neighbouring tiles of a view run on one XCD and share its
// L2; a workgroup whose range is exhausted steals from the next XCD's range.
// Uneven image sizes and ray lengths are balanced by the queue.
//
// Per ray the work is that of the fused tiled kernel (forward march writing
// transmittance checkpoints every K samples, L1/L2 loss gradient in
// registers, reverse march recomputing K-sample segments).  A workgroup keeps
// the current TF and its gradient table in LDS across tiles and flushes the
// table (one global atomic per entry) only when the next tile uses another TF
// and at the end.  Camera gradients are carried through the ReferenceFrame
// camera (renderer_camera.cuh:46-61, its adjoint as generateRaysAdjoint,
// camera_impl.cu:302-357) into the 3x3 viewport, wave-reduced and added with
// nine atomics per wave.
#pragma once
#include "dvr_tiled_impl.hpp"

namespace dvr {

template <int TFM>
__device__ __forceinline__ void batched_flush_tf(const BatchParams& B, int tfi, const float* s_g, int RC, bool priv)
{
    float* g = B.tfs[tfi].adj;
    for (int e = threadIdx.x; e < RC; e += kBatchBlock) {
        float v;
        if (priv) {
            v = 0.0f;
            for (int t = 0; t < kBatchBlock; ++t) v += s_g[e * kBatchBlock + t];
        } else if (B.tfCopyStride > 0) {
            v = 0.0f;
#pragma unroll
            for (int k = 0; k < kTfCopies; ++k) v += s_g[k * B.tfCopyStride + e];
        } else {
            v = s_g[e];
        }
        if (v != 0.0f) gatomic_add(g, unsigned(e), v);
    }
}

// Pops the next tile for this workgroup (thread 0 only): own XCD's range
// first, then the following ones.  Returns -1 when every range is exhausted.
__device__ __forceinline__ int batched_pop(const BatchParams& B, int xcd, int& qoff)
{
    while (qoff < kNumXcd) {
        const int q = (xcd + qoff) % kNumXcd;
        const int len = B.qBegin[q + 1] - B.qBegin[q];
        if (len > 0) {
            const unsigned i = atomicAdd(&B.queues[q], 1u);
            if (i < unsigned(len)) return B.qBegin[q] + int(i);
        }
        ++qoff;
    }
    return -1;
}

template <int FILTER, int TFM, int BLEND, int K, int GM>
__global__ __launch_bounds__(kBatchBlock) __attribute__((amdgpu_waves_per_eu(
    GM == 0 ? 5 : ((GM & 2) ? batched_vol_waves(GM) : 2)))) void k_batched(BatchParams B)
{
    constexpr bool ADJ = GM != 0;
    extern __shared__ __attribute__((aligned(16))) float lds[];
    __shared__ int s_tile;
    float* s_tf = lds;
    float* s_g = lds + B.ldsTf;
    const int xcd = int(blockIdx.x) % kNumXcd;
    const int lane = threadIdx.x & 63;
    const int ww = 1 << B.waveShift;
    const int lx = lane & (ww - 1);
    const int ly = lane >> B.waveShift;
    float* ck = ADJ ? B.ck + size_t(blockIdx.x) * size_t(B.ckCap) * kBatchBlock + threadIdx.x : nullptr;
    // this lane's density store: sample i at ds[i * kBatchBlock]
    float* ds = ADJ && B.ds ? B.ds + size_t(blockIdx.x) * size_t(B.dsCap) * kBatchBlock + threadIdx.x : nullptr;
    const bool priv = B.usePriv != 0;
    // volume gradients: the wave's corner values are summed per 64-byte chunk in LDS and flushed
    // into the volume's bricked scratch at the end of every tile (each tile belongs to one view,
    // hence one volume); k_debrick adds the scratch into the caller's gradient after the launch
    constexpr bool kVolInst = (GM & 2) != 0;
    using VS = std::conditional_t<kVolInst, VolAgg<BrickLayout>, VolDirect>;
    VS vs;
    const bool aggOn = kVolInst && B.wantVol && B.ldsAgg > 0;
    if constexpr (kVolInst) {
        if (aggOn) vs.init(lds + B.ldsAgg, B.aggLog2);
    } else {
        vs.p = nullptr;
    }
    int qoff = 0;
    int curTf = -1, curRC = 0;
    bool curWantTf = false;
    float lossLocal = 0.0f;

    for (;;) {
        if (threadIdx.x == 0) s_tile = batched_pop(B, xcd, qoff);
        __syncthreads();
        const int tile = __builtin_amdgcn_readfirstlane(s_tile);
        if (tile < 0) break;
        // view of the tile: the last view whose first tile is <= tile
        int lo = 0, hi = B.nViews - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (B.views[mid].tileBegin <= tile) lo = mid;
            else hi = mid - 1;
        }
        const BView& V = B.views[lo];
        const int tfi = V.tf;
        if (tfi != curTf) {  // uniform over the workgroup
            if (ADJ && curWantTf) {
                __syncthreads();
                batched_flush_tf<TFM>(B, curTf, s_g, curRC, priv);
            }
            __syncthreads();
            const BTf tf = B.tfs[tfi];
            curTf = tfi;
            curRC = tf.R * tf.C;
            curWantTf = ADJ && B.wantTf && tf.adj != nullptr;
            // rows at a stride of tf_lds_stride<TFM>() floats (16-byte aligned RGBA for the linear TF)
            for (int e = threadIdx.x; e < curRC; e += kBatchBlock)
                s_tf[(e / tf_cols<TFM>()) * tf_lds_stride<TFM>() + (e % tf_cols<TFM>())] = gld(tf.p, unsigned(e));
            if (curWantTf) {
                const int ng = priv ? curRC * kBatchBlock : (B.tfCopyStride > 0 ? kTfCopies * B.tfCopyStride : curRC);
                for (int e = threadIdx.x; e < ng; e += kBatchBlock) s_g[e] = 0.0f;
            }
            __syncthreads();
        }
        const BVol vol = B.vols[V.vol];
        if constexpr (kVolInst) {
            // every lane binds (uniform): the flush hands table entries to all lanes, also to
            // lanes outside the image
            if (aggOn && vol.adj != nullptr) {
                vs.bind(vol.brick);
                vs.L.set(vol.X, vol.Y, vol.Z);
            }
        }
        const int tl = tile - V.tileBegin;
        const int ty = tl / V.tilesX, tx = tl - ty * V.tilesX;
        const int x = tx * B.tileW + lx, y = ty * B.tileH + ly;
        const bool valid = x < V.W && y < V.H;
        constexpr int kC = tf_cols<TFM>();  // validated: every TF of the plan has this many columns
        const Tf T{s_tf, curRC / kC, kC, tf_lds_stride<TFM>(), 1};
        SinkTiled<TFM> sink;
        sink.init(s_g, kC, priv, kBatchBlock, priv ? 0 : B.tfCopyStride);
        const bool wantCam = ADJ && (GM & 4) && B.wantCam && V.acam != nullptr;
        float cv[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) cv[k] = 0.0f;
        if (valid) {
            d3 start, dir;
            d3 eye = dk3(0, 0, 0), right = eye, up = eye;
            if (V.camMode == 1) {
                double m[16];
#pragma unroll
                for (int k = 0; k < 16; ++k) m[k] = gld(V.cam, unsigned(k));
                camera_inverse_view(m, x, y, V.W, V.H, start, dir);
            } else {
                eye = dk3(gld(V.cam, 0), gld(V.cam, 1), gld(V.cam, 2));
                right = dk3(gld(V.cam, 3), gld(V.cam, 4), gld(V.cam, 5));
                up = dk3(gld(V.cam, 6), gld(V.cam, 7), gld(V.cam, 8));
                camera_reference_frame(eye, right, up, V.tanX, V.tanY, x, y, V.W, V.H, start, dir);
            }
            const Ray r = make_ray(start, dir, dk3(vol.bmin[0], vol.bmin[1], vol.bmin[2]),
                                   dk3(vol.bsize[0], vol.bsize[1], vol.bsize[2]), vol.X, vol.Y, vol.Z, V.step);
            const Vol VV{vol.p, vol.X, vol.Y, vol.Z, vol.Y * vol.Z, vol.Z, 1};
            Acc acc;
            int n;
            march_forward<FILTER, TFM, BLEND>(VV, T, r, B.earlyOut, -1, acc, n, ck, kBatchBlock, K * B.ckEvery,
                                              ADJ ? B.ckCap : 0, ds, kBatchBlock, B.dsCap);
            const unsigned pix = unsigned(y * V.W + x);
            gst(V.color, 4 * pix, acc.r);
            gst(V.color, 4 * pix + 1, acc.g);
            gst(V.color, 4 * pix + 2, acc.b);
            gst(V.color, 4 * pix + 3, acc.A);
            gsti(V.term, pix, n);
            if constexpr (ADJ) {
                f4 adj_acc;
                if (B.lossKind == DVR_LOSS_NONE) {
                    adj_acc = mk4(gld(V.acol, 4 * pix), gld(V.acol, 4 * pix + 1), gld(V.acol, 4 * pix + 2),
                                  gld(V.acol, 4 * pix + 3));
                } else {
                    const float d0 = acc.r - gld(V.target, 4 * pix), d1 = acc.g - gld(V.target, 4 * pix + 1),
                                d2 = acc.b - gld(V.target, 4 * pix + 2), d3_ = acc.A - gld(V.target, 4 * pix + 3);
                    const float sc = V.lossScale;
                    if (B.lossKind == DVR_LOSS_L1) {
                        auto sgn = [](float v) { return v > 0.f ? 1.f : (v < 0.f ? -1.f : 0.f); };
                        adj_acc = mk4(sgn(d0) * sc, sgn(d1) * sc, sgn(d2) * sc, sgn(d3_) * sc);
                        lossLocal += (fabsf(d0) + fabsf(d1) + fabsf(d2) + fabsf(d3_)) * sc;
                    } else {
                        adj_acc = mk4(2.f * d0 * sc, 2.f * d1 * sc, 2.f * d2 * sc, 2.f * d3_ * sc);
                        lossLocal += (d0 * d0 + d1 * d1 + d2 * d2 + d3_ * d3_) * sc;
                    }
                }
                f3 adj_start = mk3(0.f, 0.f, 0.f), adj_dir = mk3(0.f, 0.f, 0.f);
                float adj_step = 0.0f;
                const bool wantVol = aggOn && vol.adj != nullptr;
                // the store holds every sample of every active ray of the wave (else re-interpolate;
                // the choice is uniform over the wave, so the segment loop does not diverge)
                const bool stored = ds != nullptr && __all(n <= B.dsCap);
                march_reverse<FILTER, TFM, BLEND, K, GM, SinkTiled<TFM>, true, VS>(VV, T, r, n, adj_acc, ck,
                                                                                   kBatchBlock, B.ckCap, B.ckEvery,
                                                                                   sink, curWantTf,
                                                         wantVol, wantCam, vs,
                                                         adj_start, adj_dir, adj_step, stored ? ds : nullptr,
                                                         kBatchBlock);
                if constexpr ((GM & 4) != 0) {
                    if (wantCam) {
                        // ray (eye, normalize(front + fx tanX right + fy tanY up)) -> viewport rows
                        const f3 fe = mk3(float(eye.x), float(eye.y), float(eye.z));
                        const f3 fr = mk3(float(right.x), float(right.y), float(right.z));
                        const f3 fu = mk3(float(up.x), float(up.y), float(up.z));
                        (void)fe;
                        const float tX = float(V.tanX), tY = float(V.tanY);
                        const float fx = (2.0f * (x + 0.5f) / float(V.W) - 1.0f) * tX;
                        const float fy = (2.0f * (y + 0.5f) / float(V.H) - 1.0f) * tY;
                        const f3 front = cross3(fu, fr);
                        const f3 dd = add3(add3(front, scl3(fr, fx)), scl3(fu, fy));
                        const f3 ad = adj_normalize(dd, adj_dir);
                        f3 ar = scl3(ad, fx), au = scl3(ad, fy);
                        adj_cross(fu, fr, ad, au, ar);
                        cv[0] = adj_start.x; cv[1] = adj_start.y; cv[2] = adj_start.z;
                        cv[3] = ar.x; cv[4] = ar.y; cv[5] = ar.z;
                        cv[6] = au.x; cv[7] = au.y; cv[8] = au.z;
                    }
                }
            }
        }
        if constexpr (ADJ) {
            if constexpr (kVolInst) {
                if (aggOn && vol.adj != nullptr) vs.flush();  // uniform: one view per tile
            }
            if (curWantTf) sink.template finish<TFM == 2 ? 5 : 4>();
            if constexpr ((GM & 4) != 0) {
                if (wantCam) {  // uniform: one view per tile
#pragma unroll
                    for (int k = 0; k < 9; ++k) {
                        const float s = wave_sum(cv[k]);
                        if (lane == 0 && s != 0.0f) gatomic_add(V.acam, unsigned(k), s);
                    }
                }
            }
        }
        __syncthreads();  // s_tile and the LDS tables are reused by the next tile
    }
    if constexpr (ADJ) {
        if (curWantTf) {
            __syncthreads();
            batched_flush_tf<TFM>(B, curTf, s_g, curRC, priv);
        }
        if (B.lossOut) {
            const float s = wave_sum(lossLocal);
            if (lane == 0 && s != 0.0f) atomicAdd(B.lossOut, s);
        }
    }
}

// Launch (occ == NULL) or report the resident workgroups per CU of one
// instance (occ != NULL) for LDS bytes `lds`.
template <int FILTER, int TFM, int BLEND, int K, int GM>
static hipError_t batched_k(const BatchParams& P, int grid, size_t lds, hipStream_t s, int* occ)
{
    auto kern = k_batched<FILTER, TFM, BLEND, K, GM>;
    if (occ) return hipOccupancyMaxActiveBlocksPerMultiprocessor(occ, kern, kBatchBlock, lds);
    if (grid <= 0) return hipSuccess;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kBatchBlock), lds, s, P);
    return hipGetLastError();
}

template <int FILTER, int BLEND, int K>
static hipError_t batched_fbk(const BatchParams& P, int tfm, int gm, int grid, size_t lds, hipStream_t s, int* occ)
{
#define DVR_BGM_CASES(TV)                                                          \
    switch (gm) {                                                                  \
    case 0:                                                                        \
        if constexpr (K == 16) return batched_k<FILTER, TV, BLEND, K, 0>(P, grid, lds, s, occ); \
        else return hipErrorInvalidValue;                                          \
    case 1: return batched_k<FILTER, TV, BLEND, K, 1>(P, grid, lds, s, occ);       \
    case 3: return batched_k<FILTER, TV, BLEND, K, 3>(P, grid, lds, s, occ);       \
    case 5: return batched_k<FILTER, TV, BLEND, K, 5>(P, grid, lds, s, occ);       \
    case 7: return batched_k<FILTER, TV, BLEND, K, 7>(P, grid, lds, s, occ);       \
    default: return hipErrorInvalidValue;                                          \
    }
    switch (tfm) {
    case 0: DVR_BGM_CASES(0)
    case 1: DVR_BGM_CASES(1)
    case 2: DVR_BGM_CASES(2)
    case 3: DVR_BGM_CASES(3)
    default: return hipErrorInvalidValue;
    }
#undef DVR_BGM_CASES
}

}  // namespace dvr
