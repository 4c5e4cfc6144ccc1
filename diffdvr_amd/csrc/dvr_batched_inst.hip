// One (filter, blend, K) slice of the batched persistent kernel instances;
// the build compiles this file once per slice so the slices compile in parallel.
#include "dvr_batched_impl.hpp"

#ifndef DVR_INST_F
#error "DVR_INST_F, DVR_INST_B and DVR_INST_K must be defined"
#endif

#define DVR_CAT2(a, b) a##b
#define DVR_CAT(a, b) DVR_CAT2(a, b)
#define DVR_INST_NAME DVR_CAT(DVR_CAT(DVR_CAT(DVR_CAT(DVR_CAT(batched_inst_, DVR_INST_F), _), DVR_INST_B), _), DVR_INST_K)

namespace dvr {
hipError_t DVR_INST_NAME(const BatchParams& P, int tfm, int gm, int grid, size_t lds, hipStream_t s, int* occ)
{
    return batched_fbk<DVR_INST_F, DVR_INST_B, DVR_INST_K>(P, tfm, gm, grid, lds, s, occ);
}
}  // namespace dvr
