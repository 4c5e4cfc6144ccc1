// Tiled forward launcher and the dispatcher over the checkpointed-adjoint
// instance slices (see dvr_tiled_impl.hpp for the kernels).
#include "dvr_tiled_impl.hpp"

namespace dvr {

#define DVR_DECL_INST(F, B, K)                                                                                 \
    hipError_t adj_tiled_inst_##F##_##B##_##K(const RenderParams&, const AdjointParams&, const CkptParams&,     \
                                              const LossParams&, bool, bool, int, hipStream_t);
#define DVR_DECL_K(F, B) DVR_DECL_INST(F, B, 8) DVR_DECL_INST(F, B, 16) DVR_DECL_INST(F, B, 32)
DVR_DECL_K(0, 0)
DVR_DECL_K(0, 1)
DVR_DECL_K(1, 0)
DVR_DECL_K(1, 1)
#undef DVR_DECL_K
#undef DVR_DECL_INST

size_t occupancy_lds(size_t need, int wgPerCu)
{
    if (wgPerCu < 2) return need;
    const size_t l = size_t(kLdsPerCu) / size_t(wgPerCu + 1) + 512;
    return need > l ? need : l;
}

int tiled_num_tiles(const RenderParams& P)
{
    const int tx = (P.W + kTileW - 1) / kTileW, ty = (P.H + kTileH - 1) / kTileH;
    return tx * ty * P.B;
}

template <int FILTER, int TFM, int BLEND>
static hipError_t fwd_tiled_t(const RenderParams& P, const CkptParams& CK, hipStream_t s)
{
    const int tilesX = (P.W + kTileW - 1) / kTileW, tilesY = (P.H + kTileH - 1) / kTileH;
    const int nTiles = tilesX * tilesY * P.B;
    if (nTiles == 0) return hipSuccess;
    const size_t lds = occupancy_lds(sizeof(float) * size_t(P.R) * P.C, tiled_wg_per_cu(TFM));
    hipLaunchKernelGGL((k_forward_tiled<FILTER, TFM, BLEND>), dim3(nTiles), dim3(kBlock), lds, s, P, CK, tilesX,
                       tilesX * tilesY, nTiles);
    return hipGetLastError();
}

hipError_t launch_forward_tiled(const RenderParams& P, const CkptParams& CK, int filter, int tfm, int blend,
                                hipStream_t s)
{
    DVR_DISPATCH3(filter, tfm, blend, return fwd_tiled_t, (P, CK, s));
    return hipErrorInvalidValue;
}

hipError_t launch_adjoint_tiled(const RenderParams& P, const AdjointParams& A, const CkptParams& CK,
                                const LossParams& L, bool fused, bool ckFromFwd, int filter, int tfm, int blend,
                                hipStream_t s)
{
    if (filter < 0 || filter > 1 || blend < 0 || blend > 1) return hipErrorInvalidValue;
#define DVR_CALL(F, B, K) return adj_tiled_inst_##F##_##B##_##K(P, A, CK, L, fused, ckFromFwd, tfm, s)
#define DVR_SEL_K(F, B)                      \
    switch (CK.K) {                          \
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
