// Instantiations of the leaf GEMM kernel for double, EPI_ATOMIC_C (see leaf_gemm_kernel.hpp): single-term
// launches with affine K addressing, every register tile.
#include "leaf_dispatch.hpp"

namespace mistra {

hipError_t launch_leaf_f64_atomic(int tile, int terms, int kmode, const GemmArgs& g, bool a_mfast, bool b_nfast,
                                  hipStream_t s, dim3 grid)
{
    if (terms != 1 || (kmode != KMODE_AFFINE_VEC && kmode != KMODE_AFFINE)) return hipErrorInvalidValue;
    constexpr int TK = 16;
    if (tile_is_rect(tile))
        return kmode == KMODE_AFFINE_VEC
                   ? launch_rect<double, TK, EPI_ATOMIC_C, KMODE_AFFINE_VEC>(tile, g, a_mfast, b_nfast, s, grid)
                   : launch_rect<double, TK, EPI_ATOMIC_C, KMODE_AFFINE>(tile, g, a_mfast, b_nfast, s, grid);
    if (kmode == KMODE_AFFINE_VEC)
        return tile == TILE_64
                   ? launch_variant<double, 64, 64, TK, 1, EPI_ATOMIC_C, KMODE_AFFINE_VEC>(g, a_mfast, b_nfast, s, grid)
                   : launch_variant<double, 128, 128, TK, 1, EPI_ATOMIC_C, KMODE_AFFINE_VEC>(g, a_mfast, b_nfast, s, grid);
    return tile == TILE_64
               ? launch_variant<double, 64, 64, TK, 1, EPI_ATOMIC_C, KMODE_AFFINE>(g, a_mfast, b_nfast, s, grid)
               : launch_variant<double, 128, 128, TK, 1, EPI_ATOMIC_C, KMODE_AFFINE>(g, a_mfast, b_nfast, s, grid);
}

}  // namespace mistra
