// Host-callable launchers of the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"
#include "lowp.hpp"  // kMTileRows, kMTileCols

namespace mistra {

enum TileCfg
{
    TILE_128 = 0,     // 128x128 workgroup tile (2x2 waves of 64x64)
    TILE_64 = 1,      // 64x64 workgroup tile (2x2 waves of 32x32), for small or skinny leaves
    TILE_128x64 = 2,  // 128 rows (leaf m) x 64 columns (leaf n): 2x2 waves of 64x32, single-term launches
    TILE_64x128 = 3,  // 64 x 128: 2x2 waves of 32x64, single-term launches
    kNumTiles = 4
};

// Rows (leaf m) and columns (leaf n) of a workgroup tile; tile_extent is the side of the square tiles
// (TILE_128, TILE_64), the only ones the batched and persistent launches use.
int tile_m(int tile);
int tile_n(int tile);
int tile_extent(int tile);
int tile_kdepth(int tile, int elem_size);
bool tile_is_rect(int tile);

// Leaf GEMM over scatter vectors with fused operand sums (leaf_gemm_kernel.hpp); `terms` is the
// largest number of sub-blocks in any operand sum of the launch (1, 2 or 4).
// `kmode` (KMode): the K addressing the host checked for this launch (kfast_check in the driver).
hipError_t launch_leaf_gemm(int elem_size, int epi, int tile, int terms, int kmode, GemmArgs g, bool a_mfast,
                            bool b_nfast, hipStream_t stream);
int max_terms(const ProdDesc* prods, int nprod);

// The register tile a launch really uses (the general and some gathered-K variants only exist, or
// only run without spills, with the 64x64 tile).
int effective_tile(int tile, int kmode, int terms, bool a_mfast, bool b_nfast);

// Grouped multi-C launch (batched contractions): device tables of jobs and of their first workgroups.
struct Grouped
{
    const GemmArgs* jobs;
    const int64_t* first;
    int njobs;
    int64_t blocks;
};
hipError_t launch_leaf_grouped(int elem_size, int tile, int terms, int kmode, const Grouped& q, bool a_mfast,
                               bool b_nfast, hipStream_t stream);

// Explicit operand sums (Naive impl): for product p and its terms t,
//   out[p*rows*cols + r + c*rows] = sum_t coeff_t * X[base_t + rs[roff_t + r] + cs[coff_t + c]]
// `which` selects the A (0) or B (1) terms of each ProdDesc.
struct SumArgs
{
    const void* X;
    const int64_t* rs;
    const int64_t* cs;
    int64_t rows, cols;
    const ProdDesc* prods;
    int nprod;
    int which;
    bool rows_fast;  // true if rs has unit stride (read along rows), else read along cols
    void* out;
    int64_t zoff;  // batch offset added to X
    // Translate form (rows_fast only): every term's sub-block is the first one shifted, and `prods`
    // holds the shifts as term bases (roff / coff unused): element (r, c) of term t is
    // X[base_t + rs[r] + cs[c]].  `pair`: rows come in contiguous pairs at even element offsets, read
    // with one vector load.
    bool shifted;
    bool pair;
};
hipError_t launch_operand_sums(int elem_size, const SumArgs& a, hipStream_t stream);

// All staged sums of one operand in one pass: every source sub-block (nsrc = 4 or 16, a translate of
// the first: element (r, c) of source s at X[shift[s] + rs[r] + cs[c]]) is read once per element, and
//   out[q*rows*cols + r + c*rows] = sum_s coef[q*nsrc + s] * source_s(r, c)   for q < nout,
// skipping zero coefficients.  rs has unit stride; `pair`: rows come in contiguous pairs at even
// element offsets (one vector load per source).
constexpr int kMsumMaxCv = 28;
struct MultiSumArgs
{
    const void* X;
    int64_t zoff;
    const int64_t* rs;
    const int64_t* cs;
    int64_t rows, cols;
    int nsrc, nout;
    const double* coef;
    const int64_t* shift;
    void* out;
    bool pair;
    // 2-level sums (nsrc = 16, nout = 49) whose coefficient matrix is the Kronecker square of this
    // 7 x 4 matrix (products o, i; sub-blocks a, b of the 2 x 2 split): computed in two steps,
    // U[o][b] = sum_a coef1[o][a] src(a, b), then out[7 o + i] = sum_b coef1[i][b] U[o][b]
    const double* coef1;
    // Kernel-argument copy of coef1 when it is set, else of coef when nsrc * nout <= kMsumMaxCv (ncv
    // entries; 0: use the pointers).  Read with scalar loads: the pointer forms are vector loads after
    // the output stores, and on this target each wait for them also waits for those stores.
    int ncv;
    double cv[kMsumMaxCv];
};
hipError_t launch_operand_multisums(int elem_size, const MultiSumArgs& a, hipStream_t stream);

// Combine staged products into C sub-blocks:
//   C[c_m[roff_s + r] + c_n[coff_s + c]] = beta*C + alpha * sum_{(p,coef) in list_s} coef*M_p[r + c*ms]
constexpr int kCombineMaxProds = 49;
struct CombineTerm
{
    int prod;
    double coeff;
};
struct CombineArgs
{
    void* C;
    const void* M;
    const int64_t* c_m;
    const int64_t* c_n;
    int64_t ms, ns;
    int nsub;                   // number of C sub-blocks (4^L)
    const int64_t* sub_roff;    // [nsub]
    const int64_t* sub_coff;    // [nsub]
    const int* term_begin;      // [nsub+1] into terms
    const CombineTerm* terms;
    int nprod;
    double alpha, beta;
    int64_t nbatch;
    const int64_t* batch_c;
    // Optional one-pass form (dense != 0; nsub in {1, 4, 16}, nprod <= kCombineMaxProds, every coefficient
    // +-1): pmask[p] bit s says product p feeds sub-block s, bit 16 + s that it does with coefficient -1.
    // Every M_p is then read once and added to all its sub-blocks.
    int dense;
    unsigned pmask[kCombineMaxProds];
    // element size of M: 0 = C's; 4 with an FP64 C = products staged in FP32 (the FP8 mode)
    int m_elem;
    // M_p in tiles (m_tiled != 0, FP32 M only): kMTileRows x kMTileCols tiles, tile (i, j) of M_p at
    // (j * (m_mp / kMTileRows) + i) kMTileRows kMTileCols, rows fastest inside a tile; M_p padded to
    // m_mp x m_np (multiples of the tile), products m_mp m_np apart
    int m_tiled;
    int64_t m_mp, m_np;
};
hipError_t launch_combine(int elem_size, const CombineArgs& a, hipStream_t stream);

// Elementwise helpers.
hipError_t launch_fill_uniform(int elem_size, void* ptr, int64_t n, uint64_t seed, int scale_exp_max,
                               hipStream_t stream);
hipError_t launch_copy_bytes(void* dst, const void* src, int64_t nbytes, hipStream_t stream);
// partial sums of (x-y)^2 and y^2 into out[0..1] (device doubles, zeroed by the kernel launch itself)
hipError_t launch_rel_error(int elem_size, const void* X, const void* Y, int64_t n, double* out,
                            hipStream_t stream);
// C = beta*C (beta == 0: set to zero) over a strided tensor of up to 8 dims.
struct StridedDims
{
    int ndim;
    int64_t len[8];
    int64_t stride[8];
};
hipError_t launch_scale(int elem_size, void* C, const StridedDims& d, double beta, hipStream_t stream);
// C_t = beta*C_t (beta == 0: set to zero) over every C target t that product p writes first (ProdDesc::
// first_mask), for the launch products prods[0 .. nprod): the beta pass of the atomic multi-C form.  Targets
// are translates of the first C sub-block: element (r, c) at C + base_t + c_m[r] + c_n[c], r < ms, c < ns.
hipError_t launch_scale_first_targets(void* C, const int64_t* c_m, const int64_t* c_n, int64_t ms, int64_t ns,
                                      const ProdDesc* prods, int nprod, double beta, hipStream_t stream);
// FP32 tensors in the low-precision modes run through FP64 copies: dst[i] = src[i] over n elements, and
// back, dst[off] = float(src[off]) over the offsets of a strided tensor.
hipError_t launch_f32_to_f64(const float* src, double* dst, int64_t n, hipStream_t stream);
hipError_t launch_f64_to_f32_strided(const double* src, float* dst, const StridedDims& d, hipStream_t stream);

// General index classes (the classical fallback of src/stra_internal/3t/stra_mult.cxx:1311-1382):
// dst (dense, first index fastest over `out`) = sum over the `red` dims of src.
struct ReduceArgs
{
    StridedDims out;  // lengths and source strides of the kept dims
    StridedDims red;  // lengths and source strides of the summed dims
};
hipError_t launch_reduce(int elem_size, const void* src, void* dst, const ReduceArgs& a, hipStream_t stream);

// C = R + beta*C over every element of C; R is read with stride 0 along C's C-only dims.
struct BcastArgs
{
    StridedDims c;  // lengths and strides of C
    int64_t r_stride[8];
    double beta;
};
hipError_t launch_bcast_add(int elem_size, const void* R, void* C, const BcastArgs& a, hipStream_t stream);

}  // namespace mistra
