// Low-precision leaf products with FP64 accuracy won back (gfx950 MFMA):
//
//  * Ozaki-style BF16 slicing.  Each row i of an FP64 operand S (each column j of T) is scaled by a power
//    of two 2^e so that |x| < 1, then split into `nsl` signed integer digits d_s (|d_0| <= 256,
//    |d_s| <= 128), x = sum_s d_s 2^(-8(s+1)) up to the last digit's rounding.  A digit is an integer of
//    at most 8 significant bits, exact in BF16.  The product of two digits is an integer below 2^16,
//    so an FP32 MFMA accumulation of up to 2^24 / ((d+1) 2^16) of them (d+1 digit pairs with the same
//    weight 2^(-8(d+2)) summed together) is exact; each such chunk is then added to an FP64 accumulator.
//    Pairs with s + t >= nsl are dropped (Ozaki truncation): nsl = 2 is the hi/lo split with 3 products,
//    nsl = 5 keeps 15 products and about 40 bits.
//    (Ozaki, Ogita, Oishi, Rump, "Error-free transformations of matrix multiplication by using fast
//    routines of matrix multiplication and its applications", Numer. Algorithms 59, 2012.)
//  * Ozaki-style INT8 slicing: the same scheme with base 2^7 digits |d_s| <= 64 (exact in int8), x =
//    sum_s d_s 2^(-6-7s).  A digit product is at most 2^12 in magnitude, so the INT32 MFMA accumulation
//    of a weight group is exact over 2^31 / (NSL 2^12) K elements; groups are folded into FP64 every
//    1024 K-steps of 64.  6 digits keep about 41 bits with 21 products, on the INT8 MFMA (twice the
//    BF16 rate).
//  * FP8 E4M3 with one FP32 scale per row (column) and 128-long K block: FP32 MFMA accumulation inside a
//    block, FP64 across blocks.
//
// Operand layouts (written by the slice kernels): digits / FP8 codes in [slice][product][row][k] with k
// fastest, rows padded to a multiple of 64 and k to a multiple of 32 (FP8: 128), padding zero.  The
// MFMA operand of lane l is row (l % 16), k = 8 (l / 16) + [0, 8) -- 16 (BF16) or 8 (FP8) contiguous bytes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <type_traits>

#include "lowp.hpp"

namespace mistra {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int exp_of(double maxabs)
{
    int e = 0;
    if (maxabs > 0.0) (void)frexp(maxabs, &e);  // maxabs = f 2^e, f in [0.5, 1): |x| 2^-e < 1
    return e;
}

// ---- scale exponents ---------------------------------------------------------------------------------

// S: [q][i + k * rows] (rows fastest): one thread per (q, i)
__global__ __launch_bounds__(256) void row_exp_kernel(const double* __restrict__ S, int64_t rows, int64_t cols,
                                                      int nmat, int* __restrict__ e)
{
    const int64_t n = int64_t(nmat) * rows;
    for (int64_t x = int64_t(blockIdx.x) * 256 + threadIdx.x; x < n; x += int64_t(gridDim.x) * 256)
    {
        const int64_t q = x / rows, i = x % rows;
        const double* s = S + q * rows * cols + i;
        double m = 0.0;
        for (int64_t k = 0; k < cols; k++) m = fmax(m, fabs(s[k * rows]));
        e[x] = exp_of(m);
    }
}

// ---- Ozaki digits -----------------------------------------------------------------------------------

template <int NSL>
__device__ __forceinline__ void digits(double x, __bf16 (&d)[NSL])
{
    // x in (-1, 1): d_0 = round(x 2^8), r = x 2^8 - d_0 (exact), d_1 = round(r 2^8), ...
    double r = x;
#pragma unroll
    for (int s = 0; s < NSL; s++)
    {
        r *= 256.0;
        const double q = rint(r);
        r -= q;
        d[s] = __bf16(float(q));
    }
}

// S (rows fastest) -> out[s][q][i][k] (k fastest), through a 64 x 64 LDS transpose
template <int NSL>
__global__ __launch_bounds__(256) void slice_rows_kernel(const double* __restrict__ S, int64_t rows, int64_t cols,
                                                         const int* __restrict__ e, __bf16* __restrict__ out,
                                                         int64_t rp, int64_t kp, int nmat)
{
    __shared__ double tile[64][65];
    const int64_t tr = blockIdx.x, tk = blockIdx.y, q = blockIdx.z;
    const int tid = threadIdx.x;
    const double* s = S + q * rows * cols;
    for (int x = tid; x < 64 * 64; x += 256)
    {
        const int r = x % 64, c = x / 64;  // consecutive threads on consecutive rows: coalesced reads
        const int64_t i = tr * 64 + r, k = tk * 64 + c;
        double v = 0.0;
        if (i < rows && k < cols) v = ldexp(s[i + k * rows], -e[q * rows + i]);
        tile[r][c] = v;
    }
    __syncthreads();
    const int64_t slab = int64_t(nmat) * rp * kp;
    for (int x = tid; x < 64 * 64; x += 256)
    {
        const int c = x % 64, r = x / 64;  // consecutive threads on consecutive k: coalesced writes
        const int64_t i = tr * 64 + r, k = tk * 64 + c;
        if (i >= rp || k >= kp) continue;
        __bf16 d[NSL];
        digits<NSL>(tile[r][c], d);
        __bf16* o = out + (q * rp + i) * kp + k;
#pragma unroll
        for (int sl = 0; sl < NSL; sl++) o[sl * slab] = d[sl];
    }
}

// ---- Ozaki INT8 digits ------------------------------------------------------------------------------

template <int NSL>
__device__ __forceinline__ void digits_i8(double x, int8_t (&d)[NSL])
{
    // x in (-1, 1): r = x 2^6, d_0 = round(r) in [-64, 64], then r = (r - d) 2^7 in [-64, 64], ... (exact)
    double r = x * 64.0;
#pragma unroll
    for (int s = 0; s < NSL; s++)
    {
        const double q = rint(r);
        d[s] = int8_t(int(q));
        r = (r - q) * 128.0;
    }
}

// S (rows fastest) -> out[s][q][i][k] (k fastest, int8), through a 64 x 64 LDS transpose
template <int NSL>
__global__ __launch_bounds__(256) void slice_rows_i8_kernel(const double* __restrict__ S, int64_t rows, int64_t cols,
                                                            const int* __restrict__ e, int8_t* __restrict__ out,
                                                            int64_t rp, int64_t kp, int nmat)
{
    __shared__ double tile[64][65];
    const int64_t tr = blockIdx.x, tk = blockIdx.y, q = blockIdx.z;
    const int tid = threadIdx.x;
    const double* s = S + q * rows * cols;
    for (int x = tid; x < 64 * 64; x += 256)
    {
        const int r = x % 64, c = x / 64;
        const int64_t i = tr * 64 + r, k = tk * 64 + c;
        double v = 0.0;
        if (i < rows && k < cols) v = ldexp(s[i + k * rows], -e[q * rows + i]);
        tile[r][c] = v;
    }
    __syncthreads();
    const int64_t slab = int64_t(nmat) * rp * kp;
    for (int x = tid; x < 64 * 64; x += 256)
    {
        const int c = x % 64, r = x / 64;
        const int64_t i = tr * 64 + r, k = tk * 64 + c;
        if (i >= rp || k >= kp) continue;
        int8_t d[NSL];
        digits_i8<NSL>(tile[r][c], d);
        int8_t* o = out + (q * rp + i) * kp + k;
#pragma unroll
        for (int sl = 0; sl < NSL; sl++) o[sl * slab] = d[sl];
    }
}

// T (k fastest, [q][j][k]) -> out[s][q][j][k] (int8)
// ---- K-fast operands: scale exponent and digits in one pass -------------------------------------------

// T (k fastest, [q][j][k], kk long) -> e[q][j] and out[s][q][j][k] (rows padded to cp, k to kp, padding
// zero): one wave per padded row, which takes the row's largest magnitude (exponent e, so |x| 2^-e < 1)
// and then reads the row again for the digits -- the second read mostly from cache, where a separate
// exponent pass read the operand from HBM once more.  D = __bf16 (BF16 digits) or int8_t (INT8 digits).
template <int NSL, typename D>
__global__ __launch_bounds__(256) void slice_kfast_kernel(const double* __restrict__ T, int64_t kk, int64_t cols,
                                                          int* __restrict__ e, D* __restrict__ out, int64_t cp,
                                                          int64_t kp, int nmat)
{
    const int lane = threadIdx.x & 63;
    const int64_t n = int64_t(nmat) * cp, slab = n * kp;
    for (int64_t x = (int64_t(blockIdx.x) * 256 + threadIdx.x) / 64; x < n; x += int64_t(gridDim.x) * 4)
    {
        const int64_t q = x / cp, j = x % cp;
        const bool real = j < cols;
        const double* t = T + (q * cols + j) * kk;
        double m = 0.0;
        if (real)
            for (int64_t k = lane; k < kk; k += 64) m = fmax(m, fabs(t[k]));
        for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
        const int ex = exp_of(m);
        if (real && lane == 0) e[q * cols + j] = ex;
        D* o = out + x * kp;
        for (int64_t k = lane; k < kp; k += 64)
        {
            const double v = real && k < kk ? ldexp(t[k], -ex) : 0.0;
            D d[NSL];
            if constexpr (std::is_same<D, int8_t>::value)
                digits_i8<NSL>(v, d);
            else
                digits<NSL>(v, d);
#pragma unroll
            for (int sl = 0; sl < NSL; sl++) o[sl * slab + k] = d[sl];
        }
    }
}

// One-pass sums of a 1-level plan (7 sums of 4 translate sub-blocks, K-fast) and their digits in one
// launch (SumSliceArgs): one wave per padded row i, which computes the row's 7 sums twice from the source
// sub-blocks -- for the 7 largest magnitudes (exponents), then for the digits -- instead of staging the
// sums in HBM and reading them back.  The sums and digits are bit for bit those of the one-pass sums
// kernel followed by slice_kfast_kernel.
// In the exponent pass a lane takes 4 K positions (k0 + 64 u) per trip and issues their K-offset loads
// together, then all their source loads; the digit pass loads the next position's values, and the K
// offset of the one after, before forming the current one's digits.  With one position per trip and no
// look-ahead, each trip waited for its offset, then for its values, and a wave (one per row, 16 per CU
// at most) had one round of loads in flight; an offset load issued after a trip's digit stores would
// also wait for those stores (gfx950 counts loads and stores on one in-order counter).
// AFF: the K offsets are affine (SumSliceArgs::rs_step >= 0) and computed instead of loaded.
template <int NSL, typename D, bool AFF>
__global__ __launch_bounds__(256) void sum_slice_kfast_kernel(SumSliceArgs a)
{
    constexpr int NQ = 7, NS = 4;  // (shift[0..3] and coef[q][s] of SumSliceArgs)
    constexpr int UM = 4;  // K positions per lane and trip in the exponent pass
    const int lane = threadIdx.x & 63;
    const int64_t slab = int64_t(NQ) * a.cp * a.kp;
    D* __restrict__ out = static_cast<D*>(a.out);
    for (int64_t i = (int64_t(blockIdx.x) * 256 + threadIdx.x) / 64; i < a.cp; i += int64_t(gridDim.x) * 4)
    {
        const bool real = i < a.rows;
        const double* __restrict__ X = a.X + a.zoff + (real ? a.cs[i] : 0);
        // the source values of positions k0 + 64 u (k0 < ks); positions at or past ks load those of ks - 1
        // (in range, not used)
        auto load = [&](int64_t k0, auto& v) {
            constexpr int U = std::extent<std::remove_reference_t<decltype(v)>>::value;
            int64_t rk[U];
#pragma unroll
            for (int u = 0; u < U; u++)
            {
                const int64_t k = std::min<int64_t>(k0 + 64 * u, a.ks - 1);
                rk[u] = AFF ? a.rs0 + k * a.rs_step : a.rs[k];
            }
#pragma unroll
            for (int u = 0; u < U; u++)
#pragma unroll
                for (int s = 0; s < NS; s++) v[u][s] = X[a.shift[s] + rk[u]];
        };
        auto sums = [&](const double (&v)[NS], double (&y)[NQ]) {
#pragma unroll
            for (int q = 0; q < NQ; q++)
            {
                double acc = 0.0;
#pragma unroll
                for (int s = 0; s < NS; s++)
                {
                    const double cf = a.coef[q * NS + s];  // uniform
                    if (cf != 0.0) acc += cf * v[s];
                }
                y[q] = acc;
            }
        };
        double mx[NQ];
#pragma unroll
        for (int q = 0; q < NQ; q++) mx[q] = 0.0;
        if (real)
            for (int64_t k0 = lane; k0 < a.ks; k0 += 64 * UM)
            {
                double v[UM][NS];
                load(k0, v);
#pragma unroll
                for (int u = 0; u < UM; u++)
                {
                    const bool in = k0 + 64 * u < a.ks;
                    double y[NQ];
                    sums(v[u], y);
#pragma unroll
                    for (int q = 0; q < NQ; q++) mx[q] = fmax(mx[q], in ? fabs(y[q]) : 0.0);
                }
            }
        int ex[NQ];
#pragma unroll
        for (int q = 0; q < NQ; q++)
        {
            double m = mx[q];
            for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
            ex[q] = exp_of(m);
            if (real && lane == 0) a.e[q * a.rows + i] = ex[q];
        }
        // digit pass: position k's values were loaded in the previous trip, k + 64's go out before k's
        // digits are formed and stored, and so does the K offset of k + 128 (loaded, not computed): the
        // next trip's loads then wait for no store of this one
        auto koff = [&](int64_t k) { return AFF ? a.rs0 + k * a.rs_step : a.rs[k]; };
        double vn[NS] = {0.0, 0.0, 0.0, 0.0};
        int64_t rn = 0;  // K offset of position k + 64
        if (real && lane < a.ks)
        {
            const int64_t r0 = koff(lane);
#pragma unroll
            for (int s = 0; s < NS; s++) vn[s] = X[a.shift[s] + r0];
        }
        if (real && lane + 64 < a.ks) rn = koff(lane + 64);
        for (int64_t k = lane; k < a.kp; k += 64)
        {
            const bool in = real && k < a.ks;
            double v[NS];
#pragma unroll
            for (int s = 0; s < NS; s++) v[s] = vn[s];
            if (real && k + 64 < a.ks)
#pragma unroll
                for (int s = 0; s < NS; s++) vn[s] = X[a.shift[s] + rn];
            if (real && k + 128 < a.ks) rn = koff(k + 128);
            double y[NQ];
            if (in)
                sums(v, y);
            else
#pragma unroll
                for (int q = 0; q < NQ; q++) y[q] = 0.0;
#pragma unroll
            for (int q = 0; q < NQ; q++)
            {
                const double x = in ? ldexp(y[q], -ex[q]) : 0.0;
                D d[NSL];
                if constexpr (std::is_same<D, int8_t>::value)
                    digits_i8<NSL>(x, d);
                else
                    digits<NSL>(x, d);
                D* o = out + (int64_t(q) * a.cp + i) * a.kp + k;
#pragma unroll
                for (int sl = 0; sl < NSL; sl++) o[sl * slab] = d[sl];
            }
        }
    }
}

// The 2-level form: one wave per padded row i forms all 49 sums of the row from its 16 source values per K
// position (the one-pass Kronecker sums kernel's arithmetic), twice -- for the 49 exponents, then for the
// digits -- so every source row is read by one wave only.  Both passes load position k + 64's values and
// the K offset of k + 128 before working on position k (as the 1-level kernel's digit pass does).
template <int NSL, typename D, bool AFF>
__global__ __launch_bounds__(256) void sum_slice_kfast_kron_kernel(SumSliceArgs a)
{
    __shared__ int sex[4][49];  // the row exponents, per wave
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t slab = int64_t(49) * a.cp * a.kp;
    D* __restrict__ out = static_cast<D*>(a.out);
    for (int64_t i = (int64_t(blockIdx.x) * 256 + threadIdx.x) / 64; i < a.cp; i += int64_t(gridDim.x) * 4)
    {
        const bool real = i < a.rows;
        const double* __restrict__ X = a.X + a.zoff + (real ? a.cs[i] : 0);
        // sums 7 o .. 7 o + 6 at K position k (v: the 16 source values there)
        auto outer = [&](const double (&v)[16], int o, double (&y)[7]) {
            double u[4];
#pragma unroll
            for (int b = 0; b < 4; b++)
            {
                u[b] = 0.0;
#pragma unroll
                for (int q = 0; q < 4; q++)  // outer sub-block
                {
                    const double cf = a.coef[o * 4 + q];  // uniform
                    if (cf != 0.0) u[b] += cf * v[4 * (2 * (q / 2) + b / 2) + 2 * (q % 2) + b % 2];
                }
            }
#pragma unroll
            for (int p = 0; p < 7; p++)
            {
                double acc = 0.0;
#pragma unroll
                for (int b = 0; b < 4; b++)
                {
                    const double cf = a.coef[p * 4 + b];
                    if (cf != 0.0) acc += cf * u[b];
                }
                y[p] = acc;
            }
        };
        // look-ahead: vn holds the values of the lane's next position k, rn the K offset of k + 64
        auto koff = [&](int64_t k) { return AFF ? a.rs0 + k * a.rs_step : a.rs[k]; };
        double vn[16];
        int64_t rn = 0;
        auto start = [&]() {
#pragma unroll
            for (int s = 0; s < 16; s++) vn[s] = 0.0;
            if (real && lane < a.ks)
            {
                const int64_t r0 = koff(lane);
#pragma unroll
                for (int s = 0; s < 16; s++) vn[s] = X[a.shift[s] + r0];
            }
            if (real && lane + 64 < a.ks) rn = koff(lane + 64);
        };
        // v = position k's values; the loads of k + 64's values and k + 128's offset go out
        auto advance = [&](int64_t k, double (&v)[16]) {
#pragma unroll
            for (int s = 0; s < 16; s++) v[s] = vn[s];
            if (real && k + 64 < a.ks)
#pragma unroll
                for (int s = 0; s < 16; s++) vn[s] = X[a.shift[s] + rn];
            if (real && k + 128 < a.ks) rn = koff(k + 128);
        };
        double mx[49];
#pragma unroll
        for (int q = 0; q < 49; q++) mx[q] = 0.0;
        start();
        if (real)
            for (int64_t k = lane; k < a.ks; k += 64)
            {
                double v[16];
                advance(k, v);
#pragma unroll
                for (int o = 0; o < 7; o++)
                {
                    double y[7];
                    outer(v, o, y);
#pragma unroll
                    for (int p = 0; p < 7; p++) mx[7 * o + p] = fmax(mx[7 * o + p], fabs(y[p]));
                }
            }
#pragma unroll
        for (int q = 0; q < 49; q++)
        {
            double m = mx[q];
            for (int sh = 32; sh > 0; sh >>= 1) m = fmax(m, __shfl_xor(m, sh));
            const int ex = exp_of(m);
            if (lane == 0) sex[wave][q] = ex;
            if (real && lane == 0) a.e[q * a.rows + i] = ex;
        }
        __builtin_amdgcn_wave_barrier();
        start();
        for (int64_t k = lane; k < a.kp; k += 64)
        {
            const bool in = real && k < a.ks;
            double v[16];
            advance(k, v);
            // one outer product at a time (a loop, not unrolled: unrolled, the digits of all 49 sums stayed
            // live and spilled), and one sum of it at a time, its exponent from LDS
#pragma unroll 1
            for (int o = 0; o < 7; o++)
            {
                double y[7];
                outer(v, o, y);
#pragma unroll 1
                for (int p = 0; p < 7; p++)
                {
                    double yp = y[0];
#pragma unroll
                    for (int j = 1; j < 7; j++) yp = p == j ? y[j] : yp;
                    const int q = 7 * o + p;
                    const double x = in ? ldexp(yp, -sex[wave][q]) : 0.0;
                    D d[NSL];
                    if constexpr (std::is_same<D, int8_t>::value)
                        digits_i8<NSL>(x, d);
                    else
                        digits<NSL>(x, d);
                    D* op = out + (int64_t(q) * a.cp + i) * a.kp + k;
#pragma unroll
                    for (int sl = 0; sl < NSL; sl++) op[sl * slab] = d[sl];
                }
            }
        }
        __builtin_amdgcn_wave_barrier();  // the next row's exponents overwrite sex[wave]
    }
}

// ---- epilogue -----------------------------------------------------------------------------------------

// Element (row, col) of product q: staged into M_q[row + col * ms]; or, with the direct epilogue, C =
// alpha * v + beta * C at the scattered element.
__device__ __forceinline__ void put_out(const LowpGemmArgs& g, int q, double* Mq, int64_t row, int64_t col, double v)
{
    (void)q;
    if (g.C)
    {
        double* c = g.C + g.c_m[row] + g.c_n[col];
        *c = g.beta == 0.0 ? g.alpha * v : g.alpha * v + g.beta * *c;
    }
    else
        Mq[row + col * g.ms] = v;
}

// Output of a 128 x 64 tile held by 8 waves (4 x 2, 32 x 32 each, element v of a lane at row
// 4 (lane / 16) + v, column lane % 16 of its 16 x 16 block), through LDS in two halves of 64 rows:
// the waves holding a half write it to `stage` (64 columns x 64 doubles, row r of column c at position
// (r + c) % 64, which spreads both the writes and the reads over the banks), then all 512 threads write
// it out with consecutive threads on consecutive rows (C's unit-stride dimension after the plan's
// orientation, and M's), applying the row / column scale exponents.  `stage` (32 KiB) aliases the
// K-loop's LDS buffers, which are free after the loop.
constexpr int kStageBytes = 64 * 64 * 8;
// The scale exponents and (direct epilogue) C offsets of the elements a thread writes out: rows
// row0 + r and row0 + 64 + r, columns col0 + cg + 8 e.  Loaded before the K loop, whose MFMAs hide their
// latency: at a short K (32 steps at k = 1024) the epilogue's chain of index, exponent and C accesses
// after the loop was a large share of each tile's time, with one workgroup per CU to overlap nothing.
struct TileOutIdx
{
    int er[2], ec[8];
    int64_t cm[2], cn[8];
};
__device__ __forceinline__ TileOutIdx tile_out_idx(const LowpGemmArgs& g, int q, int64_t row0, int64_t col0,
                                                   const int* __restrict__ ea, const int* __restrict__ eb)
{
    TileOutIdx x;
    const int r = threadIdx.x & 63, cg = threadIdx.x >> 6;
#pragma unroll
    for (int h = 0; h < 2; h++)
    {
        const int64_t row = row0 + 64 * h + r;
        const bool in = row < g.ms;
        x.er[h] = in ? ea[int64_t(q) * g.ms + row] : 0;
        x.cm[h] = in && g.C ? g.c_m[row] : 0;
    }
#pragma unroll
    for (int e = 0; e < 8; e++)
    {
        const int64_t col = col0 + cg + 8 * e;
        const bool in = col < g.ns;
        x.ec[e] = in ? eb[int64_t(q) * g.ns + col] : 0;
        x.cn[e] = in && g.C ? g.c_n[col] : 0;
    }
    return x;
}

__device__ __forceinline__ void tile_out_128x64(const LowpGemmArgs& g, int q, const double (&acc)[2][2][4],
                                                double* stage, int64_t row0, int64_t col0, int wm, int wn,
                                                const TileOutIdx& ix, double* __restrict__ M)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double* Mq = M + int64_t(q) * g.ms * g.ns;
    const int r = tid & 63, cg = tid >> 6;
#pragma unroll
    for (int h = 0; h < 2; h++)
    {
        __syncthreads();  // the K loop (or the previous half) is done with the LDS
        if (((wave & 3) >> 1) == h)
        {
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++)
#pragma unroll
                    for (int v = 0; v < 4; v++)
                    {
                        const int c = wn + j * 16 + (lane & 15), rr = wm - 64 * h + i * 16 + 4 * (lane >> 4) + v;
                        stage[c * 64 + ((rr + c) & 63)] = acc[i][j][v];
                    }
        }
        __syncthreads();
        const int64_t row = row0 + 64 * h + r;
        if (row < g.ms)
        {
#pragma unroll
            for (int e = 0; e < 8; e++)
            {
                const int c = cg + 8 * e;
                const int64_t col = col0 + c;
                if (col >= g.ns) continue;
                const double v = ldexp(stage[c * 64 + ((r + c) & 63)], ix.er[h] + ix.ec[e]);
                if (g.C)
                {
                    double* cp = g.C + ix.cm[h] + ix.cn[e];
                    *cp = g.beta == 0.0 ? g.alpha * v : g.alpha * v + g.beta * *cp;
                }
                else
                    Mq[row + col * g.ms] = v;
            }
        }
    }
}

// ---- Ozaki BF16 GEMM --------------------------------------------------------------------------------

// Digit pairs (s, t) with s + t = d share the weight 2^(-8(d+2)) and are summed in one FP32 accumulator,
// which stays exact while every partial sum is an integer of magnitude <= 2^24.  With |d_0| <= 256 and
// |d_s| <= 128 (s >= 1), a K element adds at most 256^2 = 2^16 to group 0 and at most
// 2 * 256 * 128 + (d - 1) * 128^2 = (d + 3) 2^14 to group d >= 1, so group 0 holds 256 K elements and
// group d >= 1 holds 1024 / (d + 3): flushed to FP64 every 8 K-steps of 32 (d = 0, 1), 6 (d = 2),
// 5 (d = 3), 4 (d = 4, 5), 3 (d = 6).
__host__ __device__ constexpr int flush_period(int d)
{
    return d == 0 ? 8 : 32 / (d + 3);
}

// One 64 x 64 tile of M_q per workgroup (4 waves, 32 x 32 each as 2 x 2 MFMA blocks of 16 x 16).
template <int NSL>
__global__ __launch_bounds__(256) void ozaki_gemm_kernel(const __bf16* __restrict__ As, const __bf16* __restrict__ Bs,
                                                         const int* __restrict__ ea, const int* __restrict__ eb,
                                                         double* __restrict__ M, LowpGemmArgs g)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tiles_m = int(g.mp / 64);
    const int tm = blockIdx.x % tiles_m, tn = blockIdx.x / tiles_m;
    const int q = g.q0 + int(blockIdx.y);  // product (the launch covers [q0, q0 + nlaunch))
    const int64_t row0 = int64_t(tm) * 64 + (wave & 1) * 32, col0 = int64_t(tn) * 64 + (wave >> 1) * 32;
    const int64_t slab_a = int64_t(g.nprod) * g.mp * g.kp, slab_b = int64_t(g.nprod) * g.np * g.kp;
    // lane's operand rows / columns and k offset
    const __bf16* pa = As + (int64_t(q) * g.mp + row0 + (lane & 15)) * g.kp + 8 * (lane >> 4);
    const __bf16* pb = Bs + (int64_t(q) * g.np + col0 + (lane & 15)) * g.kp + 8 * (lane >> 4);

    double acc[2][2][4];
    f32x4 a32[NSL][2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
        {
#pragma unroll
            for (int v = 0; v < 4; v++) acc[i][j][v] = 0.0;
#pragma unroll
            for (int d = 0; d < NSL; d++) a32[d][i][j] = f32x4{0, 0, 0, 0};
        }
    auto flush = [&](int d) {
        const double w = ldexp(1.0, -8 * (d + 2));
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
            {
#pragma unroll
                for (int v = 0; v < 4; v++) acc[i][j][v] += w * double(a32[d][i][j][v]);
                a32[d][i][j] = f32x4{0, 0, 0, 0};
            }
    };

    const int64_t nsteps = g.kp / 32;
    for (int64_t st = 0; st < nsteps; st++)
    {
        bf16x8 fa[NSL][2], fb[NSL][2];
#pragma unroll
        for (int s = 0; s < NSL; s++)
#pragma unroll
            for (int i = 0; i < 2; i++)
            {
                fa[s][i] = *reinterpret_cast<const bf16x8*>(pa + s * slab_a + int64_t(i) * 16 * g.kp + st * 32);
                fb[s][i] = *reinterpret_cast<const bf16x8*>(pb + s * slab_b + int64_t(i) * 16 * g.kp + st * 32);
            }
#pragma unroll
        for (int s = 0; s < NSL; s++)
#pragma unroll
            for (int t = 0; s + t < NSL; t++)
#pragma unroll
                for (int i = 0; i < 2; i++)
#pragma unroll
                    for (int j = 0; j < 2; j++)
                        a32[s + t][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[s][i], fb[t][j], a32[s + t][i][j], 0, 0, 0);
#pragma unroll
        for (int d = 0; d < NSL; d++)
            if ((st + 1) % flush_period(d) == 0) flush(d);
    }
#pragma unroll
    for (int d = 0; d < NSL; d++) flush(d);

    // M_q[row + col * ms] = acc * 2^(ea[row] + eb[col]); element v of a lane: row 4 (l / 16) + v, col l % 16
    double* Mq = M + int64_t(q) * g.ms * g.ns;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
        {
            const int64_t col = col0 + j * 16 + (lane & 15);
            if (col >= g.ns) continue;
            const int ec = eb[int64_t(q) * g.ns + col];
#pragma unroll
            for (int v = 0; v < 4; v++)
            {
                const int64_t row = row0 + i * 16 + 4 * (lane >> 4) + v;
                if (row < g.ms) put_out(g, q, Mq, row, col, ldexp(acc[i][j][v], ea[int64_t(q) * g.ms + row] + ec));
            }
        }
}

constexpr uint32_t kBufOob = 0x80000000u;  // buffer-load offsets at or past this read 0

// LDS-staged form: 512 threads (8 waves as 4 x 2, 32 x 32 each) per 128 x 64 tile of M_q.  Every digit
// element of a K-step (32) is loaded from HBM/L2 once per workgroup straight into LDS (buffer loads with
// the LDS destination: no staging registers), laid out as [slice][row][64 bytes] with the four 16-byte
// chunks of row r XOR-swizzled by (r >> 1) & 3, the INT8 GEMM's layout: the MFMA operand reads (16
// consecutive rows at one chunk) spread each 8 rows over the 8 chunk slots of a 128-byte bank window
// (rows padded to 80 bytes instead cost twice the LDS cycles in bank conflicts).  A tiles are loaded two
// K-steps ahead into three buffers, B tiles one step ahead into two (5 slices: 120 + 40 KiB, all of the
// LDS): the 8 waves of the one workgroup per CU meet at every step's barrier together, so a step's MFMAs
// must cover the loads' latency by themselves.  Needs mp % 128 == 0, np % 64 == 0, kp % 32 == 0, NSL <= 5.
// Two workgroups per CU need at most 128 VGPRs: asked for with 2 slices (the 3-product split).
template <int NSL>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(NSL <= 2 ? 4 : 1))) void ozaki_gemm_lds_kernel(const __bf16* __restrict__ As, const __bf16* __restrict__ Bs,
                                                             const int* __restrict__ ea, const int* __restrict__ eb,
                                                             double* __restrict__ M, LowpGemmArgs g)
{
    static_assert(NSL >= 2 && NSL <= 5, "LDS: 3 x NSL x 8 KiB (A) + 2 x NSL x 4 KiB (B) <= 160 KiB");
    // separate arrays per buffer, indexed by compile-time buffer numbers: the compiler's wait insertion
    // then knows that a load into one buffer does not feed reads of another (with one array and a
    // run-time index it waits for every load into LDS before the next step's reads)
    // (the first A buffer, at least 4 slices long, doubles as the output staging tile; 2 slices take
    // 80 KiB, two workgroups per CU)
    __shared__ __attribute__((aligned(16))) char sa0[NSL < 4 ? 4 : NSL][128 * 64];
    __shared__ __attribute__((aligned(16))) char sa1[NSL][128 * 64], sa2[NSL][128 * 64];
    __shared__ __attribute__((aligned(16))) char sb0[NSL][64 * 64], sb1[NSL][64 * 64];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int tiles_m = int(g.mp / 128);
    const int tm = blockIdx.x % tiles_m, tn = blockIdx.x / tiles_m;
    const int q = g.q0 + int(blockIdx.y);  // product (the launch covers [q0, q0 + nlaunch))
    const int64_t row0 = int64_t(tm) * 128, col0 = int64_t(tn) * 64;
    // the epilogue's exponents and C offsets, loaded here so that the K loop hides their latency (not with
    // 2 slices, whose 128-register budget for four waves per SIMD they would overflow)
    constexpr bool kEarlyIdx = NSL > 2;
    TileOutIdx oix;
    if constexpr (kEarlyIdx) oix = tile_out_idx(g, q, row0, col0, ea, eb);
    const int wm = (wave & 3) * 32, wn = (wave >> 2) * 32;
    const int64_t slab_a = int64_t(g.nprod) * g.mp * g.kp, slab_b = int64_t(g.nprod) * g.np * g.kp;
    // Loads: a scalar resource per slice at the tile's first row, a fixed lane offset and the K-step as
    // the scalar offset (no vector instructions on addresses in the loop).  Wave w fills rows
    // 16 w .. 16 w + 15 of every A slice tile (waves 0-3 also of every B slice tile), one 16-byte chunk
    // per lane: LDS slot `lane` of the wave's 1 KiB is row 16 w + lane / 4 at position lane % 4, which
    // holds chunk (lane % 4) ^ ((row >> 1) & 3) of the row.  Byte offsets stay below 256 kp < 2^31
    // (host-checked).
    __amdgpu_buffer_rsrc_t rsa[NSL], rsb[NSL];
#pragma unroll
    for (int s = 0; s < NSL; s++)
    {
        rsa[s] = __builtin_amdgcn_make_buffer_rsrc(const_cast<__bf16*>(As + s * slab_a + (int64_t(q) * g.mp + row0) * g.kp),
                                                   (short)0, int(kBufOob), 0x00020000);
        rsb[s] = __builtin_amdgcn_make_buffer_rsrc(const_cast<__bf16*>(Bs + s * slab_b + (int64_t(q) * g.np + col0) * g.kp),
                                                   (short)0, int(kBufOob), 0x00020000);
    }
    const int dr = wave * 16 + (lane >> 2);
    const int dch = (lane & 3) ^ ((dr >> 1) & 3);
    const uint32_t vo = uint32_t((dr * g.kp + dch * 8) * 2);  // the B tile's rows are those of waves 0-3

    double acc[2][2][4];
    f32x4 a32[NSL][2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
        {
#pragma unroll
            for (int v = 0; v < 4; v++) acc[i][j][v] = 0.0;
#pragma unroll
            for (int d = 0; d < NSL; d++) a32[d][i][j] = f32x4{0, 0, 0, 0};
        }
    auto flush = [&](int d) {
        const double w = ldexp(1.0, -8 * (d + 2));
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
            {
#pragma unroll
                for (int v = 0; v < 4; v++) acc[i][j][v] += w * double(a32[d][i][j][v]);
                a32[d][i][j] = f32x4{0, 0, 0, 0};
            }
    };

    typedef __attribute__((address_space(3))) void* lds_ptr;
    auto dma_a = [&](char (*buf)[128 * 64], int st) {
#if defined(__HIP_DEVICE_COMPILE__)  // the builtin exists for the gfx950 target only, not the host pass
#pragma unroll
        for (int s = 0; s < NSL; s++)
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rsa[s], (lds_ptr)(&buf[s][wave * 1024]), 16, vo, st * 64, 0, 0);
#else
        (void)buf;
        (void)st;
#endif
    };
    auto dma_b = [&](char (*buf)[64 * 64], int st) {
#if defined(__HIP_DEVICE_COMPILE__)
        if (wave < 4)
#pragma unroll
            for (int s = 0; s < NSL; s++)
                __builtin_amdgcn_raw_ptr_buffer_load_lds(rsb[s], (lds_ptr)(&buf[s][wave * 1024]), 16, vo, st * 64, 0, 0);
#else
        (void)buf;
        (void)st;
#endif
    };
    // Before a barrier that publishes a step's tiles: wait for this wave's loads into LDS (vmcnt counts
    // them) except, with `keep`, the NSL newest ones (A of two steps ahead, issued after everything the
    // barrier publishes).  Written out: the compiler's wait insertion does not see LDS reads depending on
    // earlier loads into LDS.
    auto landed = [&](bool keep) {
#if defined(__HIP_DEVICE_COMPILE__)
        if (keep)
            __builtin_amdgcn_s_waitcnt(0x0F70 | NSL);  // vmcnt(NSL); expcnt and lgkmcnt unconstrained
        else
            __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
#else
        (void)keep;
#endif
    };

    // operand reads: 16 consecutive rows (lane % 16) at chunk lane / 16
    const int fr = lane & 15, fc = lane >> 4;
    int aoff[2], boff[2];
#pragma unroll
    for (int i = 0; i < 2; i++)
    {
        const int ra_ = wm + i * 16 + fr, rb_ = wn + i * 16 + fr;
        aoff[i] = ra_ * 64 + ((fc ^ ((ra_ >> 1) & 3)) * 16);
        boff[i] = rb_ * 64 + ((fc ^ ((rb_ >> 1) & 3)) * 16);
    }

    // The step barrier: the wave's LDS reads done (their fragments are consumed by then anyway), then
    // s_barrier, with no memory fence: __syncthreads() would also wait for every load into LDS, the A
    // tiles two steps ahead included.  The "memory" clobber keeps the compiler from moving LDS accesses
    // across it.
    auto step_barrier = [&]() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); };

    const int nsteps = int(g.kp / 32);
    // One K-step.  A of step st is in buffer st % 3 (sa0, sa1, sa2), B in buffer st % 2 (sb0, sb1).
    auto step = [&](int st, auto abuf_c) {
        constexpr int ab = decltype(abuf_c)::value;
        const int bb = st & 1;
        char(*const ca)[128 * 64] = ab == 0 ? sa0 : ab == 1 ? sa1 : sa2;
        char(*const cb)[64 * 64] = bb == 0 ? sb0 : sb1;
        bf16x8 fa[NSL][2], fb[NSL][2];
#pragma unroll
        for (int s = 0; s < NSL; s++)
#pragma unroll
            for (int i = 0; i < 2; i++)
            {
                fa[s][i] = *reinterpret_cast<const bf16x8*>(&ca[s][aoff[i]]);
                fb[s][i] = *reinterpret_cast<const bf16x8*>(&cb[s][boff[i]]);
            }
        // B of step st + 1 into the other B buffer and A of step st + 2 into the A buffer after next: both
        // last read in step st - 1, before the barrier that ended it
        __builtin_amdgcn_sched_barrier(0);
        if (st + 1 < nsteps) dma_b(bb == 0 ? sb1 : sb0, st + 1);
        if (st + 2 < nsteps) dma_a(ab == 0 ? sa2 : ab == 1 ? sa0 : sa1, st + 2);
#pragma unroll
        for (int s = 0; s < NSL; s++)
#pragma unroll
            for (int t = 0; s + t < NSL; t++)
#pragma unroll
                for (int i = 0; i < 2; i++)
#pragma unroll
                    for (int j = 0; j < 2; j++)
                        a32[s + t][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[s][i], fb[t][j], a32[s + t][i][j], 0, 0, 0);
#pragma unroll
        for (int d = 0; d < NSL; d++)
            if ((st + 1) % flush_period(d) == 0) flush(d);
        landed(st + 2 < nsteps);
        step_barrier();
    };
    typedef std::integral_constant<int, 0> I0;
    typedef std::integral_constant<int, 1> I1;
    typedef std::integral_constant<int, 2> I2;

    if (nsteps > 0)  // (no K: no loads; the tile is written as zeros)
    {
        dma_a(sa0, 0);
        dma_b(sb0, 0);
    }
    if (nsteps > 1) dma_a(sa1, 1);
    landed(nsteps > 1);
    step_barrier();
    // three steps per trip, one per A buffer
    for (int st = 0; st < nsteps; st += 3)
    {
        step(st, I0());
        if (st + 1 >= nsteps) break;
        step(st + 1, I1());
        if (st + 2 >= nsteps) break;
        step(st + 2, I2());
    }
#pragma unroll
    for (int d = 0; d < NSL; d++) flush(d);

    static_assert(sizeof(sa0) >= kStageBytes, "output staging fits the first A buffer");
    if constexpr (!kEarlyIdx) oix = tile_out_idx(g, q, row0, col0, ea, eb);
    tile_out_128x64(g, q, acc, reinterpret_cast<double*>(&sa0[0][0]), row0, col0, wm, wn, oix, M);
}

// ---- Ozaki INT8 GEMM ---------------------------------------------------------------------------------

// 512 threads (8 waves as 4 x 2, 32 x 32 each) per 128 x 64 tile of M_q, K-steps of 64.  The digit tiles
// are double-buffered in LDS as [slice][row][64 bytes] with the four 16-byte chunks of row r
// XOR-swizzled by (r >> 1) & 3, so the MFMA operand reads (16 consecutive rows at one chunk) spread each
// 8 rows over the 8 chunk slots of a 128-byte bank window.  Group d (digit pairs s + t = d, weight
// 2^(-12-7d)) accumulates exactly in INT32; with FOLD (kp > 64 kI8Fold) it is folded into FP64 every
// kI8Fold K-steps, else once at the end (which keeps the FP64 accumulators out of the loop's registers).
// Needs mp % 128 == 0, np % 64 == 0, kp % 64 == 0, NSL <= 6 (LDS).
typedef int i32x4 __attribute__((ext_vector_type(4)));
constexpr int kI8Fold = 1024;  // K-steps per INT32 fold: 1024 * 64 * NSL * 2^12 < 2^31
template <int NSL, bool FOLD>
__global__ __launch_bounds__(512) void ozaki_i8_gemm_kernel(const int8_t* __restrict__ As, const int8_t* __restrict__ Bs,
                                                            const int* __restrict__ ea, const int* __restrict__ eb,
                                                            double* __restrict__ M, LowpGemmArgs g)
{
    static_assert(NSL <= 6, "LDS: 2 x NSL x 12 KiB");
    __shared__ __attribute__((aligned(16))) int8_t sa[2][NSL][128 * 64];  // (>= 2 slices: the 32 KiB output stage)
    __shared__ __attribute__((aligned(16))) int8_t sb[2][NSL][64 * 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int tiles_m = int(g.mp / 128);
    const int tm = blockIdx.x % tiles_m, tn = blockIdx.x / tiles_m;
    const int q = g.q0 + int(blockIdx.y);  // product (the launch covers [q0, q0 + nlaunch))
    const int64_t row0 = int64_t(tm) * 128, col0 = int64_t(tn) * 64;
    // the epilogue's exponents and C offsets, loaded here so that the K loop hides their latency (not in
    // the folding form, which has no registers left for them)
    constexpr bool kEarlyIdx = !FOLD;
    TileOutIdx oix;
    if constexpr (kEarlyIdx) oix = tile_out_idx(g, q, row0, col0, ea, eb);
    const int wm = (wave & 3) * 32, wn = (wave >> 2) * 32;
    const int64_t slab_a = int64_t(g.nprod) * g.mp * g.kp, slab_b = int64_t(g.nprod) * g.np * g.kp;
    // loader: A tile = 128 rows x 64 bytes = 512 x 16 B per slice (one per thread); B = 64 x 64 = 256 x 16 B.
    // Buffer loads: a scalar resource per slice at the tile's first row, a fixed 32-bit lane offset and the
    // K-step as the scalar offset, so the loop spends no vector instructions on addresses (the vector unit
    // does not run beside the MFMAs); lanes past the B tile read 0 through an out-of-range offset.
    // Offsets stay below 128 kp < 2^31 (host-checked).
    const int lr = tid >> 2, lc = tid & 3;
    const bool bload = tid < 256;
    __amdgpu_buffer_rsrc_t rsa[NSL], rsb[NSL];
#pragma unroll
    for (int s = 0; s < NSL; s++)
    {
        rsa[s] = __builtin_amdgcn_make_buffer_rsrc(const_cast<int8_t*>(As + s * slab_a + (int64_t(q) * g.mp + row0) * g.kp),
                                                   (short)0, int(kBufOob), 0x00020000);
        rsb[s] = __builtin_amdgcn_make_buffer_rsrc(const_cast<int8_t*>(Bs + s * slab_b + (int64_t(q) * g.np + col0) * g.kp),
                                                   (short)0, int(kBufOob), 0x00020000);
    }
    const uint32_t voa = uint32_t(lr * g.kp + lc * 16);
    const uint32_t vob = bload ? uint32_t((lr & 63) * g.kp + lc * 16) : kBufOob;
    const int soff = lr * 64 + ((lc ^ ((lr >> 1) & 3)) * 16);  // this thread's swizzled LDS chunk

    double acc[2][2][4];
    i32x4 a32[NSL][2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
        {
#pragma unroll
            for (int v = 0; v < 4; v++) acc[i][j][v] = 0.0;
#pragma unroll
            for (int d = 0; d < NSL; d++) a32[d][i][j] = i32x4{0, 0, 0, 0};
        }
    auto fold = [&]() {
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
#pragma unroll
                for (int v = 0; v < 4; v++)
                {
                    // small weights first
                    double t = 0.0;
#pragma unroll
                    for (int d = NSL - 1; d >= 0; d--) t += ldexp(double(a32[d][i][j][v]), -12 - 7 * d);
                    acc[i][j][v] += t;
                }
#pragma unroll
        for (int d = 0; d < NSL; d++)
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++) a32[d][i][j] = i32x4{0, 0, 0, 0};
    };

    i32x4 ra[NSL], rb[NSL];
    auto gload = [&](int st) {
#pragma unroll
        for (int s = 0; s < NSL; s++)
        {
            ra[s] = __builtin_amdgcn_raw_buffer_load_b128(rsa[s], voa, st * 64, 0);
            rb[s] = __builtin_amdgcn_raw_buffer_load_b128(rsb[s], vob, st * 64, 0);
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int s = 0; s < NSL; s++)
        {
            *reinterpret_cast<i32x4*>(&sa[buf][s][soff]) = ra[s];
            if (bload) *reinterpret_cast<i32x4*>(&sb[buf][s][soff]) = rb[s];
        }
    };

    // operand reads: 16 consecutive rows (lane % 16) at chunk lane / 16
    const int fr = lane & 15, fc = lane >> 4;
    int aoff[2], boff[2];
#pragma unroll
    for (int i = 0; i < 2; i++)
    {
        const int ra_ = wm + i * 16 + fr, rb_ = wn + i * 16 + fr;
        aoff[i] = ra_ * 64 + ((fc ^ ((ra_ >> 1) & 3)) * 16);
        boff[i] = rb_ * 64 + ((fc ^ ((rb_ >> 1) & 3)) * 16);
    }

    const int nsteps = int(g.kp / 64);
    if (nsteps > 0)  // (no K: no loads; the tile is written as zeros)
    {
        gload(0);
        lstore(0);
    }
    __syncthreads();
    for (int st = 0; st < nsteps; st++)
    {
        const int cur = st & 1;
        if (st + 1 < nsteps) gload(st + 1);
        i32x4 fa[NSL][2];
#pragma unroll
        for (int s = 0; s < NSL; s++)
#pragma unroll
            for (int i = 0; i < 2; i++) fa[s][i] = *reinterpret_cast<const i32x4*>(&sa[cur][s][aoff[i]]);
#pragma unroll
        for (int t = 0; t < NSL; t++)
        {
            i32x4 fb[2];
#pragma unroll
            for (int j = 0; j < 2; j++) fb[j] = *reinterpret_cast<const i32x4*>(&sb[cur][t][boff[j]]);
#pragma unroll
            for (int s = 0; s + t < NSL; s++)
#pragma unroll
                for (int i = 0; i < 2; i++)
#pragma unroll
                    for (int j = 0; j < 2; j++)
                        a32[s + t][i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[s][i], fb[j], a32[s + t][i][j], 0, 0, 0);
        }
        if constexpr (FOLD)
            if ((st + 1) % g.fold_steps == 0) fold();
        if (st + 1 < nsteps) lstore(cur ^ 1);
        __syncthreads();
    }
    fold();

    static_assert(sizeof(sa) >= kStageBytes, "output staging fits the A buffers");
    if constexpr (!kEarlyIdx) oix = tile_out_idx(g, q, row0, col0, ea, eb);
    tile_out_128x64(g, q, acc, reinterpret_cast<double*>(&sa[0][0][0]), row0, col0, wm, wn, oix, M);
}

// ---- FP8 E4M3 with per-128 K-block power-of-two scales (MX-style) ------------------------------------

// Codes: E4M3 values of x / 2^e with 2^e the smallest power of two that brings the block's largest
// magnitude to at most 240 (the largest finite E4M3 value common to the OCP and FNUZ encodings), one e
// per (row, 128-long K block).  Scales: E8M0 bytes (e + 127) packed so that the four 16-row blocks of a
// 64-row band share one 32-bit word: byte (r % 64) / 16 of word [(q * nkb + kb) * (rp / 4) + (r / 64) * 16
// + r % 16].  A lane of the GEMM then loads one word per operand and K-step, and the MFMA's operand
// select picks the byte of its 16-row block.
constexpr double kE4M3Max = 240.0;

__device__ __forceinline__ uint32_t fp8x4(float a, float b, float c, float d)
{
    int lo = __builtin_amdgcn_cvt_pk_fp8_f32(a, b, 0, false);
    return uint32_t(__builtin_amdgcn_cvt_pk_fp8_f32(c, d, lo, true));
}

// exponent e of the block scale: amax / 2^e <= 240 with e as small as possible, e in [-127, 127]
__device__ __forceinline__ int fp8_block_exp(double amax)
{
    if (!(amax > 0.0)) return 0;
    int x = 0;
    const double f = frexp(amax / kE4M3Max, &x);  // amax / 240 = f 2^x, f in [0.5, 1)
    const int e = f == 0.5 ? x - 1 : x;
    return e < -127 ? -127 : (e > 127 ? 127 : e);
}

__device__ __forceinline__ int64_t fp8_scale_byte(int64_t q, int64_t kb, int64_t nkb, int64_t r, int64_t rp)
{
    return ((q * nkb + kb) * (rp / 4) + (r / 64) * 16 + (r % 16)) * 4 + (r % 64) / 16;
}

// S (rows fastest) -> codes[q][i][k] (k fastest, kp a multiple of 128) and scale bytes: one thread per
// (q, i, kb) (reads strided by `rows`, coalesced across threads; writes 128 contiguous bytes)
__global__ __launch_bounds__(256) void fp8_rows_kernel(const double* __restrict__ S, int64_t rows, int64_t cols,
                                                       uint8_t* __restrict__ out, uint8_t* __restrict__ sc,
                                                       int64_t rp, int64_t kp, int nmat)
{
    const int64_t nkb = kp / 128, n = int64_t(nmat) * rp * nkb;
    for (int64_t x = int64_t(blockIdx.x) * 256 + threadIdx.x; x < n; x += int64_t(gridDim.x) * 256)
    {
        const int64_t i = x % rp, kb = (x / rp) % nkb, q = x / (rp * nkb);
        const double* s = S + q * rows * cols + i;
        double m = 0.0;
        if (i < rows)
            for (int64_t k = kb * 128; k < kb * 128 + 128 && k < cols; k++) m = fmax(m, fabs(s[k * rows]));
        const int e = fp8_block_exp(m);
        sc[fp8_scale_byte(q, kb, nkb, i, rp)] = uint8_t(e + 127);
        uint32_t* o = reinterpret_cast<uint32_t*>(out + (q * rp + i) * kp + kb * 128);
        for (int c = 0; c < 128; c += 4)
        {
            float v[4];
            for (int u = 0; u < 4; u++)
            {
                const int64_t k = kb * 128 + c + u;
                v[u] = (i < rows && k < cols) ? float(ldexp(s[k * rows], -e)) : 0.0f;
            }
            o[c / 4] = fp8x4(v[0], v[1], v[2], v[3]);
        }
    }
}

// T (k fastest, [q][j][k]) -> codes[q][j][k] and scale bytes: one wave per (q, j, kb)
__global__ __launch_bounds__(256) void fp8_cols_kernel(const double* __restrict__ T, int64_t kk, int64_t cols,
                                                       uint8_t* __restrict__ out, uint8_t* __restrict__ sc,
                                                       int64_t cp, int64_t kp, int nmat)
{
    const int lane = threadIdx.x & 63;
    const int64_t nkb = kp / 128, n = int64_t(nmat) * cp * nkb;
    for (int64_t x = (int64_t(blockIdx.x) * 256 + threadIdx.x) / 64; x < n; x += int64_t(gridDim.x) * 4)
    {
        const int64_t kb = x % nkb, j = (x / nkb) % cp, q = x / (nkb * cp);
        const int64_t k0 = kb * 128 + 2 * lane;
        const double* t = T + (q * cols + j) * kk;
        const double v0 = (j < cols && k0 < kk) ? t[k0] : 0.0, v1 = (j < cols && k0 + 1 < kk) ? t[k0 + 1] : 0.0;
        double m = fmax(fabs(v0), fabs(v1));
        for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
        const int e = fp8_block_exp(m);
        if (lane == 0) sc[fp8_scale_byte(q, kb, nkb, j, cp)] = uint8_t(e + 127);
        const int pk = __builtin_amdgcn_cvt_pk_fp8_f32(float(ldexp(v0, -e)), float(ldexp(v1, -e)), 0, false);
        reinterpret_cast<uint16_t*>(out + (q * cp + j) * kp + kb * 128)[lane] = uint16_t(pk & 0xffff);
    }
}

// 512 threads (8 waves as 4 x 2, 64 x 64 each = 4 x 4 blocks of 16 x 16) per 256 x 128 tile of M_q,
// K-steps of 128 (one v_mfma_scale_f32_16x16x128_f8f6f4 per block and step, the block scales applied by
// the MFMA itself, FP32 accumulation across blocks).  The code tiles and their scale words go from HBM/L2
// straight into LDS (buffer loads with the LDS destination, no staging registers), three K-steps deep:
// the loads of step st + 2 are issued while step st computes (three 50 KiB stages: one workgroup per CU,
// two waves per SIMD).  No load in the loop writes registers, so the compiler adds no waits of its own
// for loads.  LDS holds each tile as [row][128 bytes] with the eight 16-byte chunks of row r XOR-swizzled
// by r & 7, so the operand reads (16 consecutive rows at one chunk pair) spread over the chunk slots.
// Products staged in FP32 (m_f32) leave straight from the accumulators into M_q's 256 x 128 tile layout
// (CombineArgs::m_tiled): a tile is 128 KiB of contiguous memory, in which a lane holds 4 consecutive
// rows of one column per block, one 16-byte store each.  The C update of a single product goes through
// LDS.  Needs mp % 256 == 0, np % 128 == 0, kp % 128 == 0, 256 kp < 2^31 and (kp / 128) max(mp, np) <
// 2^31 (host-checked).
typedef int i32x8 __attribute__((ext_vector_type(8)));
typedef int i32x4v __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(512) void fp8_gemm_kernel(const uint8_t* __restrict__ As, const uint8_t* __restrict__ Bs,
                                                      const uint32_t* __restrict__ sa, const uint32_t* __restrict__ sb,
                                                      double* __restrict__ M, LowpGemmArgs g)
{
    // One array per stage, indexed by compile-time stage numbers: the A tile (32 KiB), the B tile (16 KiB),
    // then 256 bytes of scale words per wave (wave 0: the tile's 64 A words, wave 1: its 32 B words and 32
    // zeros, the others zeros).  Three arrays keep the compiler's tracking of loads into LDS exact (with
    // more, it lost track and waited for loads into other stages before every read).
    constexpr int kB = 256 * 128, kW = kB + 128 * 128, kStage = kW + 8 * 256;
    static_assert(kMTileRows == 256 && kMTileCols == 128, "the FP8 GEMM tile is M's tile");
    __shared__ __attribute__((aligned(16))) uint8_t ls0[kStage], ls1[kStage], ls2[kStage];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = (wave & 3) * 64, wn = (wave >> 2) * 64;
    const int64_t tiles_m = g.mp / 256;
    const int tm = int(blockIdx.x % tiles_m), tn = int(blockIdx.x / tiles_m);
    const int q = g.q0 + int(blockIdx.y);  // product (the launch covers [q0, q0 + nlaunch))
    const int64_t row0 = int64_t(tm) * 256, col0 = int64_t(tn) * 128;
    const int64_t nkb = g.kp / 128;
    const int nsteps = int(nkb);

    // Loads: scalar resources at the tile's first row (code tiles) and first scale word, fixed lane
    // offsets and the K-step as the scalar offset.  Wave w fills A rows 32 w .. 32 w + 31 (four 1 KiB
    // loads) and B rows 16 w .. 16 w + 15 (two): LDS slot `lane` of a 1 KiB load is row 8 u + lane / 8 of
    // the load's block at position lane % 8, which holds chunk (lane % 8) ^ (lane / 8) of the row (row & 7
    // = lane / 8).  Scale words: those of rows row0 .. row0 + 255 are 64 consecutive words of a K-step's
    // block (4 groups of 16 rows), those of columns col0 .. col0 + 127 32 words.  Wave 0 loads the A words
    // of a step, wave 1 the B words (lanes past 32 read 0), the other waves read 0: one load per wave and
    // step into the wave's own slot, no branch.
    const __amdgpu_buffer_rsrc_t rsa = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<uint8_t*>(As + (int64_t(q) * g.mp + row0) * g.kp), (short)0, int(kBufOob), 0x00020000);
    const __amdgpu_buffer_rsrc_t rsb = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<uint8_t*>(Bs + (int64_t(q) * g.np + col0) * g.kp), (short)0, int(kBufOob), 0x00020000);
    const uint32_t* w = wave == 1 ? sb + int64_t(q) * nkb * (g.np / 4) + (col0 / 64) * 16
                                  : sa + int64_t(q) * nkb * (g.mp / 4) + (row0 / 64) * 16;
    const __amdgpu_buffer_rsrc_t rw =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t*>(w), (short)0, int(kBufOob), 0x00020000);
    const int lrow = lane >> 3, lch = (lane & 7) ^ (lane >> 3);
    uint32_t voa[4], vob[2];
#pragma unroll
    for (int u = 0; u < 4; u++) voa[u] = uint32_t((32 * wave + 8 * u + lrow) * g.kp + lch * 16);
#pragma unroll
    for (int u = 0; u < 2; u++) vob[u] = uint32_t((16 * wave + 8 * u + lrow) * g.kp + lch * 16);
    const uint32_t vow = wave == 0 || (wave == 1 && lane < 32) ? uint32_t(lane * 4) : kBufOob;
    const int64_t wstep = wave == 1 ? g.np : g.mp;  // bytes of scale words per K-step
    // this lane's scale words: rows wm + 16 i + lane % 16 (byte i), at (wm / 64) 16 + lane % 16 of the
    // step's 64, columns the same with wn
    const int wsel = (wm / 64) * 16 + (lane & 15), wselb = (wn / 64) * 16 + (lane & 15);
    // operand reads: row (16-row block base + lane % 16), chunks 2 (lane / 16) and 2 (lane / 16) + 1
    const int fr = lane & 15, fg = lane >> 4;
    int aoff[4][2], boff[4][2];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int h = 0; h < 2; h++)
        {
            const int ra = wm + 16 * i + fr, rb = wn + 16 * i + fr;
            aoff[i][h] = ra * 128 + (((2 * fg + h) ^ (ra & 7)) << 4);
            boff[i][h] = rb * 128 + (((2 * fg + h) ^ (rb & 7)) << 4);
        }

    f32x4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = f32x4{0, 0, 0, 0};

    typedef __attribute__((address_space(3))) void* lds_ptr;
    // The seven loads into LDS of K-step kb (the wait below counts them).  Past the last step (kb >=
    // nsteps) every offset is out of range: the loads read zeros into a stage that no later step reads,
    // so every step issues seven loads and waits the same way, with no branch inside a step.
    auto issue = [&](uint8_t* buf, int kb) {
        __builtin_amdgcn_sched_barrier(0);
#if defined(__HIP_DEVICE_COMPILE__)  // the builtin exists for the gfx950 target only, not the host pass
        const bool real = kb < nsteps;
        // (the offsets are uniform; said explicitly, or the compiler's induction variable for one can land
        // in a vector register and the load becomes a loop over its distinct values)
        const int so = __builtin_amdgcn_readfirstlane(real ? kb * 128 : 0);
        const int sw = __builtin_amdgcn_readfirstlane(real ? int(kb * wstep) : 0);
#pragma unroll
        for (int u = 0; u < 4; u++)
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rsa, (lds_ptr)(buf + (4 * wave + u) * 1024), 16,
                                                     real ? voa[u] : kBufOob, so, 0, 0);
#pragma unroll
        for (int u = 0; u < 2; u++)
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rsb, (lds_ptr)(buf + kB + (2 * wave + u) * 1024), 16,
                                                     real ? vob[u] : kBufOob, so, 0, 0);
        __builtin_amdgcn_raw_ptr_buffer_load_lds(rw, (lds_ptr)(buf + kW + wave * 256), 4, real ? vow : kBufOob, sw, 0,
                                                 0);
#else
        (void)buf;
        (void)kb;
#endif
        __builtin_amdgcn_sched_barrier(0);
    };
    // Before the barrier that publishes the next step's tiles: this wave's loads of that step done; the
    // newest 7 (those of the step after, issued after everything the barrier publishes) may stay in
    // flight.  The compiler's wait insertion does not see LDS reads depending on loads into LDS.
    auto landed = [&]() {
#if defined(__HIP_DEVICE_COMPILE__)
        __builtin_amdgcn_s_waitcnt(0x0F70 | 7);  // vmcnt(7); expcnt and lgkmcnt unconstrained
#endif
    };
    // this wave's LDS reads done, then s_barrier with no memory fence (the loads of the step after next
    // stay in flight); the "memory" clobber keeps LDS accesses on their side of it
    auto step_barrier = [&]() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); };

    // One K-step on stage S (tiles and scale words in ls<S>), loading step st + 2.  A step is one basic
    // block: with a branch between the MFMAs and the barrier, the compiler sank the MFMAs to their next
    // use, a step later, and the fragments of several steps spilled.
    auto step = [&](int st, auto stage_c) {
        constexpr int S = decltype(stage_c)::value;
        constexpr int N2 = (S + 2) % 3;  // stage of step st + 2 (last read in step st - 1)
        const uint8_t* ca = S == 0 ? ls0 : S == 1 ? ls1 : ls2;
        const uint8_t* cb = ca + kB;
        const int ws_a = reinterpret_cast<const int*>(ca + kW)[wsel];
        const int ws_b = reinterpret_cast<const int*>(ca + kW + 256)[wselb];
        i32x8 fa[4], fb[4];
#pragma unroll
        for (int i = 0; i < 4; i++)
        {
            const i32x4v b0 = *reinterpret_cast<const i32x4v*>(&cb[boff[i][0]]);
            const i32x4v b1 = *reinterpret_cast<const i32x4v*>(&cb[boff[i][1]]);
            fb[i] = i32x8{b0[0], b0[1], b0[2], b0[3], b1[0], b1[1], b1[2], b1[3]};
        }
#pragma unroll
        for (int i = 0; i < 4; i++)
        {
            const i32x4v a0 = *reinterpret_cast<const i32x4v*>(&ca[aoff[i][0]]);
            const i32x4v a1 = *reinterpret_cast<const i32x4v*>(&ca[aoff[i][1]]);
            fa[i] = i32x8{a0[0], a0[1], a0[2], a0[3], a1[0], a1[1], a1[2], a1[3]};
        }
        issue(N2 == 0 ? ls0 : N2 == 1 ? ls1 : ls2, st + 2);
#define MISTRA_FP8_MFMA(I, J) \
    acc[I][J] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(fa[I], fb[J], acc[I][J], 0, 0, I, ws_a, J, ws_b)
#define MISTRA_FP8_ROW(I) \
    MISTRA_FP8_MFMA(I, 0); \
    MISTRA_FP8_MFMA(I, 1); \
    MISTRA_FP8_MFMA(I, 2); \
    MISTRA_FP8_MFMA(I, 3)
        MISTRA_FP8_ROW(0);
        MISTRA_FP8_ROW(1);
        MISTRA_FP8_ROW(2);
        MISTRA_FP8_ROW(3);
#undef MISTRA_FP8_ROW
#undef MISTRA_FP8_MFMA
        __builtin_amdgcn_sched_barrier(0);
        landed();
        step_barrier();
        __builtin_amdgcn_sched_barrier(0);
    };
    typedef std::integral_constant<int, 0> I0;
    typedef std::integral_constant<int, 1> I1;
    typedef std::integral_constant<int, 2> I2;

    issue(ls0, 0);  // (no K: only out-of-range loads; the tile is written as zeros)
    issue(ls1, 1);
    landed();
    step_barrier();
    // three steps per trip, one per stage, then the last nsteps % 3 (no exit inside a trip: with one, the
    // compiler waited for every load into LDS at the loop's head)
    int st = 0;
    for (; st + 3 <= nsteps; st += 3)
    {
        step(st, I0());
        step(st + 1, I1());
        step(st + 2, I2());
    }
    if (st < nsteps) step(st, I0());
    if (st + 1 < nsteps) step(st + 1, I1());
    // the loads of the two steps past the end may still be writing zeros into LDS: none may outlive the
    // workgroup (its LDS goes to the next one)
#if defined(__HIP_DEVICE_COMPILE__)
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
#endif

    if (!g.C)
    {
        // into the tile of M_q (padded rows and columns included): element (r, c) of the tile at
        // c kMTileRows + r, block (i, j) of the lane at its base + 64 i + 16 kMTileRows 4 j bytes
        float* mt = reinterpret_cast<float*>(M) + int64_t(q) * g.mp * g.np + (int64_t(tn) * tiles_m + tm) * (256 * 128);
        const __amdgpu_buffer_rsrc_t rm = __builtin_amdgcn_make_buffer_rsrc(mt, (short)0, 256 * 128 * 4, 0x00020000);
        const uint32_t mbase = uint32_t(((wn + (lane & 15)) * 256 + wm + 4 * (lane >> 4)) * 4);
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++)
                __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(i32x4v, acc[i][j]), rm, mbase,
                                                       64 * i + 16384 * j, 0);
        return;
    }
    // The C update of a single product: through this wave's 16 KiB of the idle tile buffers, as
    // [column][row] with row r of column c at r ^ (c % 32), then one row per lane and column.  The
    // barrier: every wave's loads into LDS done (each waited for its own above).
    __syncthreads();
    float* stg = reinterpret_cast<float*>((wave < 3 ? ls0 : wave < 6 ? ls1 : ls2) + (wave % 3) * 16384);
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int v = 0; v < 4; v++)
            {
                const int c = 16 * j + (lane & 15), r = 16 * i + 4 * (lane >> 4) + v;
                stg[c * 64 + (r ^ (c & 31))] = acc[i][j][v];
            }
    double* Mq = M + int64_t(q) * g.ms * g.ns;
    const int64_t row = row0 + wm + lane;
#pragma unroll 1
    for (int c = 0; c < 64; c++)
    {
        const int64_t col = col0 + wn + c;
        if (col >= g.ns || row >= g.ms) continue;
        put_out(g, q, Mq, row, col, double(stg[c * 64 + (lane ^ (c & 31))]));
    }
}

unsigned grid_of(int64_t n, int per)
{
    int64_t b = (n + per - 1) / per;
    return unsigned(b < 1 ? 1 : (b > 65536 ? 65536 : b));
}

template <int NSL>
hipError_t slices_t(const LowpSliceArgs& a, hipStream_t s)
{
    if (a.skip_a)
        ;
    else if (a.s_kfast)  // S laid out as T: rows of contiguous K
        hipLaunchKernelGGL((slice_kfast_kernel<NSL, __bf16>), dim3(grid_of(int64_t(a.nprod) * a.mp, 4)), dim3(256), 0,
                           s, a.S, a.ks, a.ms, a.ea, static_cast<__bf16*>(a.outA), a.mp, a.kp, a.nprod);
    else
        hipLaunchKernelGGL(slice_rows_kernel<NSL>, dim3(unsigned(a.mp / 64), unsigned(a.kp / 64), unsigned(a.nprod)),
                           dim3(256), 0, s, a.S, a.ms, a.ks, a.ea, static_cast<__bf16*>(a.outA), a.mp, a.kp, a.nprod);
    if (!a.skip_b)
        hipLaunchKernelGGL((slice_kfast_kernel<NSL, __bf16>), dim3(grid_of(int64_t(a.nprod) * a.np, 4)), dim3(256), 0, s,
                           a.T, a.ks, a.ns, a.eb, static_cast<__bf16*>(a.outB), a.np, a.kp, a.nprod);
    return hipGetLastError();
}

template <int NSL>
hipError_t gemm_t(const void* As, const void* Bs, const int* ea, const int* eb, double* M, const LowpGemmArgs& g,
                  hipStream_t s)
{
    if constexpr (NSL <= 5)  // the double-buffered digit tiles of 6 or 7 slices exceed the 160 KiB of LDS
    {
        // (tile byte offsets below 256 kp < 2^31 for its buffer loads; longer K takes the other kernel)
        if (g.mp % 128 == 0 && g.np % 64 == 0 && g.kp % 32 == 0 && g.kp < (int64_t(1) << 23))
        {
            hipLaunchKernelGGL(ozaki_gemm_lds_kernel<NSL>, dim3(unsigned((g.mp / 128) * (g.np / 64)), unsigned(g.nlaunch)),
                               dim3(512), 0, s, static_cast<const __bf16*>(As), static_cast<const __bf16*>(Bs), ea, eb,
                               M, g);
            return hipGetLastError();
        }
    }
        hipLaunchKernelGGL(ozaki_gemm_kernel<NSL>, dim3(unsigned((g.mp / 64) * (g.np / 64)), unsigned(g.nlaunch)),
                           dim3(256), 0, s, static_cast<const __bf16*>(As), static_cast<const __bf16*>(Bs), ea, eb, M, g);
    return hipGetLastError();
}

template <int NSL>
hipError_t slices_i8_t(const LowpSliceArgs& a, hipStream_t s)
{
    if (a.skip_a)
        ;
    else if (a.s_kfast)
        hipLaunchKernelGGL((slice_kfast_kernel<NSL, int8_t>), dim3(grid_of(int64_t(a.nprod) * a.mp, 4)), dim3(256), 0,
                           s, a.S, a.ks, a.ms, a.ea, static_cast<int8_t*>(a.outA), a.mp, a.kp, a.nprod);
    else
        hipLaunchKernelGGL(slice_rows_i8_kernel<NSL>, dim3(unsigned(a.mp / 64), unsigned(a.kp / 64), unsigned(a.nprod)),
                           dim3(256), 0, s, a.S, a.ms, a.ks, a.ea, static_cast<int8_t*>(a.outA), a.mp, a.kp, a.nprod);
    if (!a.skip_b)
        hipLaunchKernelGGL((slice_kfast_kernel<NSL, int8_t>), dim3(grid_of(int64_t(a.nprod) * a.np, 4)), dim3(256), 0, s,
                           a.T, a.ks, a.ns, a.eb, static_cast<int8_t*>(a.outB), a.np, a.kp, a.nprod);
    return hipGetLastError();
}

template <int NSL>
hipError_t gemm_i8_t(const void* As, const void* Bs, const int* ea, const int* eb, double* M, const LowpGemmArgs& g,
                     hipStream_t s)
{
    if (g.mp % 128 || g.np % 64 || g.kp % 64) return hipErrorInvalidValue;
    if (g.kp >= (int64_t(1) << 24)) return hipErrorInvalidValue;  // tile offsets below 128 kp < 2^31
    const dim3 grid(unsigned((g.mp / 128) * (g.np / 64)), unsigned(g.nlaunch));
    if (g.fold_steps < 0 || g.fold_steps > kI8Fold) return hipErrorInvalidValue;
    if (g.fold_steps > 0 || g.kp > int64_t(64) * kI8Fold)
    {
        LowpGemmArgs gf = g;
        if (gf.fold_steps == 0) gf.fold_steps = kI8Fold;
        hipLaunchKernelGGL((ozaki_i8_gemm_kernel<NSL, true>), grid, dim3(512), 0, s, static_cast<const int8_t*>(As),
                           static_cast<const int8_t*>(Bs), ea, eb, M, gf);
    }
    else
        hipLaunchKernelGGL((ozaki_i8_gemm_kernel<NSL, false>), grid, dim3(512), 0, s, static_cast<const int8_t*>(As),
                           static_cast<const int8_t*>(Bs), ea, eb, M, g);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_ozaki_i8_slices(int nsl, const LowpSliceArgs& a, hipStream_t s)
{
    switch (nsl)
    {
    case 2: return slices_i8_t<2>(a, s);
    case 3: return slices_i8_t<3>(a, s);
    case 4: return slices_i8_t<4>(a, s);
    case 5: return slices_i8_t<5>(a, s);
    case 6: return slices_i8_t<6>(a, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_ozaki_i8_gemm(int nsl, const void* As, const void* Bs, const int* ea, const int* eb, double* M,
                                const LowpGemmArgs& g, hipStream_t s)
{
    switch (nsl)
    {
    case 2: return gemm_i8_t<2>(As, Bs, ea, eb, M, g, s);
    case 3: return gemm_i8_t<3>(As, Bs, ea, eb, M, g, s);
    case 4: return gemm_i8_t<4>(As, Bs, ea, eb, M, g, s);
    case 5: return gemm_i8_t<5>(As, Bs, ea, eb, M, g, s);
    case 6: return gemm_i8_t<6>(As, Bs, ea, eb, M, g, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_lowp_exponents(const LowpSliceArgs& a, hipStream_t s)
{
    // K-fast operands (T always, S when s_kfast) get their exponents in the slice pass itself
    if (!a.s_kfast && !a.skip_a)
    {
        hipLaunchKernelGGL(row_exp_kernel, dim3(grid_of(int64_t(a.nprod) * a.ms, 256)), dim3(256), 0, s, a.S, a.ms,
                           a.ks, a.nprod, a.ea);
        return hipGetLastError();
    }
    return hipSuccess;
}

template <int NSL, typename D>
hipError_t sum_slices_t(const SumSliceArgs& a, hipStream_t s)
{
    const dim3 grid(grid_of(a.cp, 4));
    if (a.level == 1 && a.rs_step >= 0) hipLaunchKernelGGL((sum_slice_kfast_kernel<NSL, D, true>), grid, dim3(256), 0, s, a);
    else if (a.level == 1) hipLaunchKernelGGL((sum_slice_kfast_kernel<NSL, D, false>), grid, dim3(256), 0, s, a);
    else if (a.level == 2 && a.rs_step >= 0)
        hipLaunchKernelGGL((sum_slice_kfast_kron_kernel<NSL, D, true>), grid, dim3(256), 0, s, a);
    else if (a.level == 2) hipLaunchKernelGGL((sum_slice_kfast_kron_kernel<NSL, D, false>), grid, dim3(256), 0, s, a);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_ozaki_sum_slices(int nsl, bool int8, const SumSliceArgs& a, hipStream_t s)
{
    if (a.cp <= 0 || a.kp <= 0) return hipSuccess;
    switch (nsl)
    {
    case 2: return int8 ? sum_slices_t<2, int8_t>(a, s) : sum_slices_t<2, __bf16>(a, s);
    case 3: return int8 ? sum_slices_t<3, int8_t>(a, s) : sum_slices_t<3, __bf16>(a, s);
    case 4: return int8 ? sum_slices_t<4, int8_t>(a, s) : sum_slices_t<4, __bf16>(a, s);
    case 5: return int8 ? sum_slices_t<5, int8_t>(a, s) : sum_slices_t<5, __bf16>(a, s);
    case 6: return int8 ? sum_slices_t<6, int8_t>(a, s) : sum_slices_t<6, __bf16>(a, s);
    case 7: return int8 ? hipErrorInvalidValue : sum_slices_t<7, __bf16>(a, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_ozaki_slices(int nsl, const LowpSliceArgs& a, hipStream_t s)
{
    switch (nsl)
    {
    case 2: return slices_t<2>(a, s);
    case 3: return slices_t<3>(a, s);
    case 4: return slices_t<4>(a, s);
    case 5: return slices_t<5>(a, s);
    case 6: return slices_t<6>(a, s);
    case 7: return slices_t<7>(a, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_ozaki_gemm(int nsl, const void* As, const void* Bs, const int* ea, const int* eb, double* M,
                             const LowpGemmArgs& g, hipStream_t s)
{
    switch (nsl)
    {
    case 2: return gemm_t<2>(As, Bs, ea, eb, M, g, s);
    case 3: return gemm_t<3>(As, Bs, ea, eb, M, g, s);
    case 4: return gemm_t<4>(As, Bs, ea, eb, M, g, s);
    case 5: return gemm_t<5>(As, Bs, ea, eb, M, g, s);
    case 6: return gemm_t<6>(As, Bs, ea, eb, M, g, s);
    case 7: return gemm_t<7>(As, Bs, ea, eb, M, g, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_fp8_slices(const LowpSliceArgs& a, uint32_t* sa, uint32_t* sb, hipStream_t s)
{
    if (a.mp % 64 || a.np % 64 || a.kp % 128) return hipErrorInvalidValue;
    hipLaunchKernelGGL(fp8_rows_kernel, dim3(grid_of(int64_t(a.nprod) * a.mp * (a.kp / 128), 256)), dim3(256), 0, s,
                       a.S, a.ms, a.ks, static_cast<uint8_t*>(a.outA), reinterpret_cast<uint8_t*>(sa), a.mp, a.kp,
                       a.nprod);
    hipLaunchKernelGGL(fp8_cols_kernel, dim3(grid_of(int64_t(a.nprod) * a.np * (a.kp / 128), 4)), dim3(256), 0, s,
                       a.T, a.ks, a.ns, static_cast<uint8_t*>(a.outB), reinterpret_cast<uint8_t*>(sb), a.np, a.kp,
                       a.nprod);
    return hipGetLastError();
}

hipError_t launch_fp8_gemm(const void* As, const void* Bs, const uint32_t* sa, const uint32_t* sb, double* M,
                           const LowpGemmArgs& g, hipStream_t s)
{
    if (g.mp % 256 || g.np % 128 || g.kp % 128 || g.kp >= (int64_t(1) << 23)) return hipErrorInvalidValue;
    if ((g.kp / 128) * std::max(g.mp, g.np) >= (int64_t(1) << 31)) return hipErrorInvalidValue;
    if (!g.C != g.m_f32) return hipErrorInvalidValue;  // products into M in FP32 tiles, or one into C
    const int64_t tiles = (g.mp / 256) * (g.np / 128);
    if (tiles > 0x7fffffff || g.nlaunch > 65535) return hipErrorInvalidValue;
    hipLaunchKernelGGL(fp8_gemm_kernel, dim3(unsigned(tiles), unsigned(g.nlaunch)), dim3(512), 0, s,
                       static_cast<const uint8_t*>(As), static_cast<const uint8_t*>(Bs), sa, sb, M, g);
    return hipGetLastError();
}

}  // namespace mistra
