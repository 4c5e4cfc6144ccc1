// Host-side launcher declarations and the mode -> template dispatch macro.
#pragma once
#include <hip/hip_runtime.h>
#include "dvr_params.hpp"

// Expands CALL<FILTER,TFM,BLEND>ARGS for the runtime (filter, tf, blend) triple.
#define DVR_DISPATCH_BLEND(F, T, blend, STMT, ARGS)        \
    switch (blend) {                                        \
    case 0: STMT<F, T, 0> ARGS; break;                      \
    case 1: STMT<F, T, 1> ARGS; break;                      \
    default: break;                                         \
    }
#define DVR_DISPATCH_TF(F, tfm, blend, STMT, ARGS)          \
    switch (tfm) {                                          \
    case 0: DVR_DISPATCH_BLEND(F, 0, blend, STMT, ARGS) break; \
    case 1: DVR_DISPATCH_BLEND(F, 1, blend, STMT, ARGS) break; \
    case 2: DVR_DISPATCH_BLEND(F, 2, blend, STMT, ARGS) break; \
    case 3: DVR_DISPATCH_BLEND(F, 3, blend, STMT, ARGS) break; \
    default: break;                                         \
    }
#define DVR_DISPATCH3(filter, tfm, blend, STMT, ARGS)       \
    switch (filter) {                                       \
    case 0: DVR_DISPATCH_TF(0, tfm, blend, STMT, ARGS) break; \
    case 1: DVR_DISPATCH_TF(1, tfm, blend, STMT, ARGS) break; \
    default: break;                                         \
    }

namespace dvr {

// Occupancy cap of the tiled kernels.  A workgroup asks for at least
// kLdsPerCu / (wgPerCu + 1) + 512 bytes of dynamic LDS, so no more than
// wgPerCu of them share a CU.  The gather-bound instances run at two 4-wave
// workgroups per CU: more rays in flight evict each other's cells from the
// 32 KB L1 faster than they hide latency (C2, 8 views 256x256 over 128^3:
// forward with checkpoints 0.93 -> 0.59 ms, checkpointed adjoint
// 0.72 -> 0.65 ms against the register-limited occupancy; volume gradients
// of 256^3 -9 %).  The Gaussian TF evaluates every component per sample and
// is arithmetic-bound: it keeps the register-limited occupancy (the cap cost
// it 3 %).
constexpr int kLdsPerCu = 160 * 1024;  // gfx950
constexpr int kTiledWgPerCu = 8;
constexpr int kTiledVolWgPerCu = 16;  // volume + TF gradient instances (four waves per SIMD)
constexpr int tiled_wg_per_cu(int tfm) { return tfm == 3 ? 0 : kTiledWgPerCu; }
size_t occupancy_lds(size_t need, int wgPerCu);

// dvr_kernels_plain.hip
hipError_t launch_forward_plain(const RenderParams& P, int filter, int tfm, int blend, hipStream_t s);
hipError_t launch_adjoint_plain(const RenderParams& P, const AdjointParams& A, int filter, int tfm, int blend,
                                hipStream_t s);

// dvr_kernels_tiled.hip
// Number of 16x16 tiles of the B x H x W image batch.
int tiled_num_tiles(const RenderParams& P);
hipError_t launch_forward_tiled(const RenderParams& P, const CkptParams& CK, int filter, int tfm, int blend,
                                hipStream_t s);
// Checkpointed adjoint.  `fused` = run the forward march in the same kernel,
// write colour/termination to P.color/P.term and form the adjoint colour from
// the loss L; otherwise A.fcol/A.fterm/A.acol are inputs.  CK.buf holds
// either checkpoints written by launch_forward_tiled (ckptFromForward) or
// scratch of tiled_num_tiles()*CK.cap*kBlock floats.
hipError_t launch_adjoint_tiled(const RenderParams& P, const AdjointParams& A, const CkptParams& CK,
                                const LossParams& L, bool fused, bool ckptFromForward, int filter, int tfm,
                                int blend, hipStream_t s);

// dvr_kernels_aux.hip
hipError_t launch_generate_rays(const float* viewport, int vSb, int vSr, int vSc, int B, int H, int W, float fovY,
                                float* rs, const int rsS[4], float* rd, const int rdS[4], hipStream_t s);
hipError_t launch_generate_rays_adjoint(const float* viewport, int vSb, int vSr, int vSc, int B, int H, int W,
                                        float fovY, const float* ars, const int arsS[4], const float* ard,
                                        const int ardS[4], float* out, int oSb, int oSr, int oSc, hipStream_t s);

// dvr_kernels_opt.hip
hipError_t launch_zero_segments(const ZeroSeg* dSegs, int nSegs, long long maxN, hipStream_t s);
// increments *dStep, then updates every block with step count *dStep
hipError_t launch_adam(const AdamSeg* dSegs, int nSegs, long long maxN, long long* dStep, double lr, double b1,
                       double b2, double eps, hipStream_t s);

// dvr_kernels_batched.hip: launch the batched persistent kernel (occ == NULL)
// or report resident workgroups per CU of that instance.  gm: 0 forward only,
// 1 TF, 3 volume (+TF), 7 volume + TF + camera.
hipError_t launch_batched(const BatchParams& P, int filter, int tfm, int blend, int K, int gm, int grid, size_t lds,
                          hipStream_t s, int* occ);
// adds every volume's bricked gradient scratch into its gradient (after launch_batched)
hipError_t launch_debrick(const DebrickSeg* dSegs, int nSegs, long long maxBricks, hipStream_t s);

// pre-shaded volumes (dvr_kernels_preshaded.hip): plain structure, inversion adjoint
hipError_t launch_forward_preshaded(const RenderParams& P, int blend, hipStream_t s);
hipError_t launch_adjoint_preshaded(const RenderParams& P, const AdjointParams& A, int blend, hipStream_t s);

// forward-mode derivatives (dvr_kernels_fwdgrad.hip)
hipError_t launch_forward_gradients(const RenderParams& P, const FwdGradParams& G, int filter, int tfm, int blend,
                                    hipStream_t s);
hipError_t launch_forward_to_adjoint(const FwdToAdjParams& F, const int* dtf, int dSb, int dSr, int dSc, float* atf,
                                     int tSb, int tSr, int tSc, int R, int C, hipStream_t s);
hipError_t launch_compare_to_image(const CompareParams& Q, hipStream_t s);

// TF utilities (dvr_kernels_tfutil.hip).  tfm 0..3 or 6 (GaussianLog, pre-shading only).
hipError_t launch_preshade(const PreshadeParams& P, int tfm, bool adjoint, hipStream_t s);
hipError_t launch_best_fit(const BestFitParams& P, int tfm, hipStream_t s);
// maxima pass (clears P.maxima first), then one decision per sample
hipError_t launch_sample_importance(const ImportanceParams& P, int tfm, hipStream_t s);

// elementwise losses (dvr_kernels_losses.hip).  op: 0 mul_log, 1 mul_log backward,
// 2 log_mse, 3 log_mse backward, 4 log_l1, 5 log_l1 backward; dbl: float64 data.
hipError_t launch_elementwise(const ElementwiseParams& P, int op, bool dbl, hipStream_t s);

}  // namespace dvr
