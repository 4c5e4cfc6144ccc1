// Dispatcher over the batched persistent kernel instance slices (kernels in
// dvr_batched_impl.hpp).
#include "dvr_launch.hpp"

namespace dvr {

#define DVR_DECL_INST(F, B, K)                                                                                  \
    hipError_t batched_inst_##F##_##B##_##K(const BatchParams&, int, int, int, size_t, hipStream_t, int*);
#define DVR_DECL_K(F, B) DVR_DECL_INST(F, B, 8) DVR_DECL_INST(F, B, 16) DVR_DECL_INST(F, B, 32)
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

}  // namespace dvr
