// Low-precision leaf products (lowp_kernels.hip): Ozaki BF16 and INT8 digit slicing and FP8 E4M3 block
// scaling.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mistra {

// Dense FP64 operands of nprod products: S[q] (ms x ks, rows fastest) and T[q] (ks x ns, k fastest);
// sliced outputs in [slice][q][row][k] (k fastest) with rows padded to mp / np (multiples of 64) and k to
// kp (a multiple of 64 for digits, 128 for FP8).
struct LowpSliceArgs
{
    const double* S;   // ms x ks rows fastest, or with s_kfast [q][i][k] (K fastest, as T)
    const double* T;
    int64_t ms, ns, ks, mp, np, kp;
    int nprod;
    int* ea;  // [q][ms] row exponents of S
    int* eb;  // [q][ns] column exponents of T
    void* outA;
    void* outB;
    bool s_kfast;
    // operands whose digits (and exponents) a fused sum-and-slice launch wrote already
    bool skip_a, skip_b;
};

// One-pass operand sums and their digits in one launch, for a 1- or 2-level plan whose operand is K-fast
// with translate sub-blocks.  At level 1, sum q (of 7) of row i at K position k is
// sum_s coef[q][s] X[zoff + shift[s] + rs[k] + cs[i]] over the 4 sub-blocks s, summed in that order with
// zero coefficients skipped (the arithmetic of the one-pass sums kernel); level 2 follows the one-pass
// Kronecker sums kernel (below).  Each row's exponent and digits then come out exactly as the K-fast
// slice kernel derives them from the staged sums.  Rows are padded to cp and K to kp with zero digits.
struct SumSliceArgs
{
    const double* X;
    int64_t zoff;
    const int64_t* rs;  // K offsets, ks of them
    int64_t rs0, rs_step;  // rs_step >= 0: rs[k] = rs0 + k rs_step (the kernel computes it), else -1
    const int64_t* cs;  // row offsets, rows of them
    int64_t ks, rows, cp, kp;
    // level 1: 4 sub-blocks and coef[q][s] for the 7 sums; level 2: 16 sub-blocks and the 7 x 4 one-level
    // coefficients coef[o][a], whose Kronecker square gives the 49 sums in the one-pass Kronecker sums
    // kernel's order (per outer product o, U_b = sum_a coef[o][a] src(a, b), then sum 7 o + p =
    // sum_b coef[p][b] U_b; source (a, b) is sub-block 4 (2 (a / 2) + b / 2) + 2 (a % 2) + b % 2)
    int level;
    int64_t shift[16];
    double coef[28];
    int* e;      // [q][rows] row exponents
    void* out;   // digits [slice][q][cp][kp]
};
hipError_t launch_ozaki_sum_slices(int nslices, bool int8, const SumSliceArgs& a, hipStream_t s);

// Tile of the FP8 products staged in FP32 (LowpGemmArgs::m_f32, CombineArgs::m_tiled)
constexpr int kMTileRows = 256, kMTileCols = 128;

struct LowpGemmArgs
{
    int64_t ms, ns, mp, np, kp;
    int nprod;  // products in the sliced operands (their slab stride)
    // direct epilogue (a single product feeding all of C with coefficient 1): C[c_m[r] + c_n[c]] =
    // alpha * M + beta * C instead of staging M
    double* C;
    const int64_t* c_m;
    const int64_t* c_n;
    double alpha, beta;
    // this launch: products [q0, q0 + nlaunch)
    int q0, nlaunch;
    // INT8 GEMM: fold the exact INT32 sums into FP64 every fold_steps K-steps of 64 (0: the default,
    // 1024, and only when K needs it; at most 1024)
    int fold_steps;
    // FP8 GEMM: stage M_q in FP32 as mp x np in 256 x 128 tiles (the layout of CombineArgs::m_tiled;
    // M is then a float array, products mp np apart)
    bool m_f32;
};

// Scale exponents, then digits: launch_lowp_exponents fills ea for an S with rows fastest; the slice
// launches fill the exponents of K-fast operands (T, and S with s_kfast) in the same pass as the digits.
hipError_t launch_lowp_exponents(const LowpSliceArgs& a, hipStream_t s);
hipError_t launch_ozaki_slices(int nslices, const LowpSliceArgs& a, hipStream_t s);
hipError_t launch_ozaki_gemm(int nslices, const void* As, const void* Bs, const int* ea, const int* eb, double* M,
                             const LowpGemmArgs& g, hipStream_t s);
hipError_t launch_ozaki_i8_slices(int nslices, const LowpSliceArgs& a, hipStream_t s);
hipError_t launch_ozaki_i8_gemm(int nslices, const void* As, const void* Bs, const int* ea, const int* eb, double* M,
                                const LowpGemmArgs& g, hipStream_t s);
// FP8: E4M3 codes and E8M0 block scales (one per row and 128-long K block, four 16-row blocks per word);
// the GEMM needs mp, np and kp multiples of 128.
hipError_t launch_fp8_slices(const LowpSliceArgs& a, uint32_t* sa, uint32_t* sb, hipStream_t s);
hipError_t launch_fp8_gemm(const void* As, const void* Bs, const uint32_t* sa, const uint32_t* sb, double* M,
                           const LowpGemmArgs& g, hipStream_t s);

}  // namespace mistra
