// Instantiations of the leaf GEMM kernel for float, EPI_STORE (see leaf_gemm_kernel.hpp).
#include "leaf_dispatch.hpp"

namespace mistra {

hipError_t launch_leaf_f32_store(int tile, int terms, int kmode, const GemmArgs& g, bool a_mfast, bool b_nfast,
                               hipStream_t s, dim3 grid)
{
    return launch_leaf_typed<float, EPI_STORE>(tile, terms, kmode, g, a_mfast, b_nfast, s, grid);
}

}  // namespace mistra
