// Launch-time dispatch of the leaf GEMM kernel template over its compile-time variants.
#pragma once

#include "leaf_gemm_kernel.hpp"

namespace mistra {

template <typename T, int TM, int TN, int TK, int TERMS, int EPI, int KF>
hipError_t launch_variant(const GemmArgs& g, bool a_mfast, bool b_nfast, hipStream_t s, dim3 grid)
{
    if (a_mfast && b_nfast)
        hipLaunchKernelGGL((leaf_gemm_kernel<T, TM, TN, TK, TERMS, true, true, EPI, KF>), grid, dim3(256), 0, s, g);
    else if (a_mfast)
        hipLaunchKernelGGL((leaf_gemm_kernel<T, TM, TN, TK, TERMS, true, false, EPI, KF>), grid, dim3(256), 0, s, g);
    else if (b_nfast)
        hipLaunchKernelGGL((leaf_gemm_kernel<T, TM, TN, TK, TERMS, false, true, EPI, KF>), grid, dim3(256), 0, s, g);
    else
        hipLaunchKernelGGL((leaf_gemm_kernel<T, TM, TN, TK, TERMS, false, false, EPI, KF>), grid, dim3(256), 0, s, g);
    return hipGetLastError();
}

// Grouped launches (batched contractions), multi-C epilogue only.

template <typename T, int TM, int TK, int TERMS, int KF>
hipError_t grouped_variant(const Grouped& q, bool a_mfast, bool b_nfast, hipStream_t s)
{
    const dim3 grid(unsigned(q.blocks));
    if (a_mfast && b_nfast)
        hipLaunchKernelGGL((leaf_gemm_grouped<T, TM, TM, TK, TERMS, true, true, EPI_MULTI_C, KF>), grid, dim3(256), 0,
                           s, q.jobs, q.first, q.njobs);
    else if (a_mfast)
        hipLaunchKernelGGL((leaf_gemm_grouped<T, TM, TM, TK, TERMS, true, false, EPI_MULTI_C, KF>), grid, dim3(256),
                           0, s, q.jobs, q.first, q.njobs);
    else if (b_nfast)
        hipLaunchKernelGGL((leaf_gemm_grouped<T, TM, TM, TK, TERMS, false, true, EPI_MULTI_C, KF>), grid, dim3(256),
                           0, s, q.jobs, q.first, q.njobs);
    else
        hipLaunchKernelGGL((leaf_gemm_grouped<T, TM, TM, TK, TERMS, false, false, EPI_MULTI_C, KF>), grid, dim3(256),
                           0, s, q.jobs, q.first, q.njobs);
    return hipGetLastError();
}

template <typename T, int TM, int TK, int KF>
hipError_t grouped_terms(const Grouped& q, int terms, bool a_mfast, bool b_nfast, hipStream_t s)
{
    if (terms <= 1) return grouped_variant<T, TM, TK, 1, KF>(q, a_mfast, b_nfast, s);
    if (terms == 2) return grouped_variant<T, TM, TK, 2, KF>(q, a_mfast, b_nfast, s);
    return grouped_variant<T, TM, TK, 4, KF>(q, a_mfast, b_nfast, s);
}

template <typename T>
hipError_t launch_grouped_typed(int tile, int terms, int kmode, const Grouped& q, bool a_mfast, bool b_nfast,
                                hipStream_t s)
{
    constexpr int TK = sizeof(T) == 8 ? 16 : 32;
    if (kmode == KMODE_GENERAL) return grouped_terms<T, 64, TK, KMODE_GENERAL>(q, terms, a_mfast, b_nfast, s);
    if (kmode == KMODE_GATHER)
        return tile == TILE_64 ? grouped_terms<T, 64, TK, KMODE_GATHER>(q, terms, a_mfast, b_nfast, s)
                               : grouped_terms<T, 128, TK, KMODE_GATHER>(q, terms, a_mfast, b_nfast, s);
    if (kmode == KMODE_AFFINE_VEC)
        return tile == TILE_64 ? grouped_terms<T, 64, TK, KMODE_AFFINE_VEC>(q, terms, a_mfast, b_nfast, s)
                               : grouped_terms<T, 128, TK, KMODE_AFFINE_VEC>(q, terms, a_mfast, b_nfast, s);
    if (tile == TILE_64) return grouped_terms<T, 64, TK, KMODE_AFFINE>(q, terms, a_mfast, b_nfast, s);
    return grouped_terms<T, 128, TK, KMODE_AFFINE>(q, terms, a_mfast, b_nfast, s);
}

template <typename T, int TM, int TK, int EPI, int KF>
hipError_t launch_terms(const GemmArgs& g, int terms, bool a_mfast, bool b_nfast, hipStream_t s, dim3 grid)
{
    if (terms <= 1) return launch_variant<T, TM, TM, TK, 1, EPI, KF>(g, a_mfast, b_nfast, s, grid);
    if (terms == 2) return launch_variant<T, TM, TM, TK, 2, EPI, KF>(g, a_mfast, b_nfast, s, grid);
    return launch_variant<T, TM, TM, TK, 4, EPI, KF>(g, a_mfast, b_nfast, s, grid);
}

// Rectangular tiles (single-term launches with affine K addressing only; effective_tile maps every other
// launch to a square tile).
template <typename T, int TK, int EPI, int KF>
hipError_t launch_rect(int tile, const GemmArgs& g, bool a_mfast, bool b_nfast, hipStream_t s, dim3 grid)
{
    if (tile == TILE_128x64) return launch_variant<T, 128, 64, TK, 1, EPI, KF>(g, a_mfast, b_nfast, s, grid);
    return launch_variant<T, 64, 128, TK, 1, EPI, KF>(g, a_mfast, b_nfast, s, grid);
}

// One element type and epilogue; `kmode` selects the K addressing (KMode).  The general (per-term
// gathered K) variant exists only with the 64x64 tile.
template <typename T, int EPI>
hipError_t launch_leaf_typed(int tile, int terms, int kmode, const GemmArgs& g, bool a_mfast, bool b_nfast,
                             hipStream_t s, dim3 grid)
{
    constexpr int TK = sizeof(T) == 8 ? 16 : 32;
    if (tile_is_rect(tile))
    {
        if (terms > 1 || (kmode != KMODE_AFFINE_VEC && kmode != KMODE_AFFINE)) return hipErrorInvalidValue;
        return kmode == KMODE_AFFINE_VEC ? launch_rect<T, TK, EPI, KMODE_AFFINE_VEC>(tile, g, a_mfast, b_nfast, s, grid)
                                         : launch_rect<T, TK, EPI, KMODE_AFFINE>(tile, g, a_mfast, b_nfast, s, grid);
    }
    if (kmode == KMODE_GENERAL) return launch_terms<T, 64, TK, EPI, KMODE_GENERAL>(g, terms, a_mfast, b_nfast, s, grid);
    if (kmode == KMODE_GATHER)
        return tile == TILE_64 ? launch_terms<T, 64, TK, EPI, KMODE_GATHER>(g, terms, a_mfast, b_nfast, s, grid)
                               : launch_terms<T, 128, TK, EPI, KMODE_GATHER>(g, terms, a_mfast, b_nfast, s, grid);
    if (kmode == KMODE_AFFINE_VEC)
        return tile == TILE_64 ? launch_terms<T, 64, TK, EPI, KMODE_AFFINE_VEC>(g, terms, a_mfast, b_nfast, s, grid)
                               : launch_terms<T, 128, TK, EPI, KMODE_AFFINE_VEC>(g, terms, a_mfast, b_nfast, s, grid);
    if (tile == TILE_64) return launch_terms<T, 64, TK, EPI, KMODE_AFFINE>(g, terms, a_mfast, b_nfast, s, grid);
    return launch_terms<T, 128, TK, EPI, KMODE_AFFINE>(g, terms, a_mfast, b_nfast, s, grid);
}

}  // namespace mistra
