// One (filter, blend, K) slice of the checkpointed adjoint instances; the
// build compiles this file once per slice so the slices compile in parallel.
#include "dvr_tiled_impl.hpp"

#ifndef DVR_INST_F
#error "DVR_INST_F, DVR_INST_B and DVR_INST_K must be defined"
#endif

#define DVR_CAT2(a, b) a##b
#define DVR_CAT(a, b) DVR_CAT2(a, b)
#define DVR_INST_NAME DVR_CAT(DVR_CAT(DVR_CAT(DVR_CAT(DVR_CAT(adj_tiled_inst_, DVR_INST_F), _), DVR_INST_B), _), DVR_INST_K)

namespace dvr {
hipError_t DVR_INST_NAME(const RenderParams& P, const AdjointParams& A, const CkptParams& CK, const LossParams& L,
                         bool fused, bool ckFromFwd, int tfm, hipStream_t s)
{
    return adj_tiled_fbk<DVR_INST_F, DVR_INST_B, DVR_INST_K>(P, A, CK, L, fused, ckFromFwd, tfm, s);
}
}  // namespace dvr
