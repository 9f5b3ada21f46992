// Host side of the leaf GEMM: grid shape and the per-type dispatch (kernels in leaf_f64.hip / leaf_f32.hip).
#include <cstdint>
#include "launch.hpp"

namespace mistra {

#define MISTRA_DECLARE_LEAF(name)                                                                     \
    hipError_t name(int tile, int terms, int kmode, const GemmArgs& g, bool a_mfast, bool b_nfast, hipStream_t s, \
                    dim3 grid);
MISTRA_DECLARE_LEAF(launch_leaf_f64_multi)
MISTRA_DECLARE_LEAF(launch_leaf_f64_store)
MISTRA_DECLARE_LEAF(launch_leaf_f32_multi)
MISTRA_DECLARE_LEAF(launch_leaf_f32_store)
MISTRA_DECLARE_LEAF(launch_leaf_f64_atomic)
#undef MISTRA_DECLARE_LEAF
hipError_t launch_leaf_f64_grouped(int tile, int terms, int kmode, const Grouped& q, bool a_mfast, bool b_nfast,
                                   hipStream_t s);
hipError_t launch_leaf_f32_grouped(int tile, int terms, int kmode, const Grouped& q, bool a_mfast, bool b_nfast,
                                   hipStream_t s);

int effective_tile(int tile, int kmode, int terms, bool a_mfast, bool b_nfast)
{
    if (kmode == KMODE_GENERAL) return TILE_64;  // the general variant exists with the 64x64 tile only
    // rectangular variants: single-term launches with affine K; a plan's other launches take 128 x 128
    if (tile_is_rect(tile) && (terms > 1 || (kmode != KMODE_AFFINE_VEC && kmode != KMODE_AFFINE))) tile = TILE_128;
    // gathered K offsets cost registers per lane: these 128x128 variants would spill
    if (kmode == KMODE_GATHER && (terms >= 4 || (terms >= 2 && a_mfast && b_nfast))) return TILE_64;
    return tile;
}

hipError_t launch_leaf_grouped(int elem_size, int tile, int terms, int kmode, const Grouped& q, bool a_mfast,
                               bool b_nfast, hipStream_t stream)
{
    if (q.njobs <= 0 || q.blocks <= 0) return hipSuccess;
    if (tile != effective_tile(tile, kmode, terms, a_mfast, b_nfast)) return hipErrorInvalidValue;  // jobs hold tile counts
    return elem_size == 8 ? launch_leaf_f64_grouped(tile, terms, kmode, q, a_mfast, b_nfast, stream)
                          : launch_leaf_f32_grouped(tile, terms, kmode, q, a_mfast, b_nfast, stream);
}

int tile_m(int tile) { return tile == TILE_64 || tile == TILE_64x128 ? 64 : 128; }
int tile_n(int tile) { return tile == TILE_64 || tile == TILE_128x64 ? 64 : 128; }
int tile_extent(int tile) { return tile == TILE_64 ? 64 : 128; }
bool tile_is_rect(int tile) { return tile == TILE_128x64 || tile == TILE_64x128; }

int tile_kdepth(int tile, int elem_size)
{
    (void)tile;
    return elem_size == 8 ? 16 : 32;
}

int max_terms(const ProdDesc* prods, int nprod)
{
    int t = 1;
    for (int p = 0; p < nprod; p++)
    {
        if (prods[p].nA > t) t = prods[p].nA;
        if (prods[p].nB > t) t = prods[p].nB;
    }
    return t;
}

hipError_t launch_leaf_gemm(int elem_size, int epi, int tile, int terms, int kmode, GemmArgs g, bool a_mfast,
                            bool b_nfast, hipStream_t stream)
{
    tile = effective_tile(tile, kmode, terms, a_mfast, b_nfast);
    const int tm = tile_m(tile), tn = tile_n(tile);
    g.tiles_m = int((g.ms + tm - 1) / tm);
    g.tiles_n = int((g.ns + tn - 1) / tn);
    if (g.tiles_m == 0 || g.tiles_n == 0 || g.nbatch == 0 || g.nprod == 0) return hipSuccess;
    if (g.ks <= 0) return hipErrorInvalidValue;  // callers route empty K to a scale of C
    if (g.ks / 16 >= int64_t(INT32_MAX)) return hipErrorInvalidValue;  // the kernels count K-tiles in 32 bits
    if (kmode == KMODE_GATHER && g.ks >= (int64_t(1) << 28)) return hipErrorInvalidValue;  // 32-bit K offsets
    // grid.y indexes the launch's products; the products of a multi-C launch must write disjoint C
    // sub-blocks (the caller groups them so), since their read-modify-write updates are not ordered
    dim3 grid(unsigned(g.tiles_m) * unsigned(g.tiles_n), unsigned(g.nprod), unsigned(g.nbatch));
    if (epi == EPI_ATOMIC_C)
        return elem_size == 8 ? launch_leaf_f64_atomic(tile, terms, kmode, g, a_mfast, b_nfast, stream, grid)
                              : hipErrorInvalidValue;
    if (elem_size == 8)
        return epi == EPI_STORE ? launch_leaf_f64_store(tile, terms, kmode, g, a_mfast, b_nfast, stream, grid)
                                : launch_leaf_f64_multi(tile, terms, kmode, g, a_mfast, b_nfast, stream, grid);
    return epi == EPI_STORE ? launch_leaf_f32_store(tile, terms, kmode, g, a_mfast, b_nfast, stream, grid)
                            : launch_leaf_f32_multi(tile, terms, kmode, g, a_mfast, b_nfast, stream, grid);
}

}  // namespace mistra
