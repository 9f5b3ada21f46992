// Launch-time dispatch of the leaf GEMM kernel template over its compile-time variants.
#pragma once

#include "leaf_gemm_kernel.hpp"

namespace mistra {

template <typename T, int TM, int TK, int TERMS, int EPI, bool KF>
hipError_t launch_variant(const GemmArgs& g, bool a_mfast, bool b_nfast, hipStream_t s, dim3 grid)
{
    if (a_mfast && b_nfast)
        hipLaunchKernelGGL((leaf_gemm_kernel<T, TM, TM, TK, TERMS, true, true, EPI, KF>), grid, dim3(256), 0, s, g);
    else if (a_mfast)
        hipLaunchKernelGGL((leaf_gemm_kernel<T, TM, TM, TK, TERMS, true, false, EPI, KF>), grid, dim3(256), 0, s, g);
    else if (b_nfast)
        hipLaunchKernelGGL((leaf_gemm_kernel<T, TM, TM, TK, TERMS, false, true, EPI, KF>), grid, dim3(256), 0, s, g);
    else
        hipLaunchKernelGGL((leaf_gemm_kernel<T, TM, TM, TK, TERMS, false, false, EPI, KF>), grid, dim3(256), 0, s, g);
    return hipGetLastError();
}

template <typename T, int TM, int TK, int EPI, bool KF>
hipError_t launch_terms(const GemmArgs& g, int terms, bool a_mfast, bool b_nfast, hipStream_t s, dim3 grid)
{
    if (terms <= 1) return launch_variant<T, TM, TK, 1, EPI, KF>(g, a_mfast, b_nfast, s, grid);
    if (terms == 2) return launch_variant<T, TM, TK, 2, EPI, KF>(g, a_mfast, b_nfast, s, grid);
    return launch_variant<T, TM, TK, 4, EPI, KF>(g, a_mfast, b_nfast, s, grid);
}

// One element type and epilogue; `kfast` selects the affine-K variant.  The general (gathered K)
// variant exists only with the 64x64 tile: it serves layouts whose K runs are not affine per tile.
template <typename T, int EPI>
hipError_t launch_leaf_typed(int tile, int terms, bool kfast, const GemmArgs& g, bool a_mfast, bool b_nfast,
                             hipStream_t s, dim3 grid)
{
    constexpr int TK = sizeof(T) == 8 ? 16 : 32;
    if (!kfast) return launch_terms<T, 64, TK, EPI, false>(g, terms, a_mfast, b_nfast, s, grid);
    if (tile == TILE_64) return launch_terms<T, 64, TK, EPI, true>(g, terms, a_mfast, b_nfast, s, grid);
    return launch_terms<T, 128, TK, EPI, true>(g, terms, a_mfast, b_nfast, s, grid);
}

}  // namespace mistra
