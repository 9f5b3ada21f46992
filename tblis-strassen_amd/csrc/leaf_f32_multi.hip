// Instantiations of the leaf GEMM kernel for float, EPI_MULTI_C (see leaf_gemm_kernel.hpp).
#include "leaf_dispatch.hpp"

namespace mistra {

hipError_t launch_leaf_f32_multi(int tile, int terms, int kmode, const GemmArgs& g, bool a_mfast, bool b_nfast,
                               hipStream_t s, dim3 grid)
{
    return launch_leaf_typed<float, EPI_MULTI_C>(tile, terms, kmode, g, a_mfast, b_nfast, s, grid);
}

hipError_t launch_leaf_f32_grouped(int tile, int terms, int kmode, const Grouped& q, bool a_mfast, bool b_nfast,
                                 hipStream_t s)
{
    return launch_grouped_typed<float>(tile, terms, kmode, q, a_mfast, b_nfast, s);
}

}  // namespace mistra
