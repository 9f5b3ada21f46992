// Leaf-product GEMM of the Strassen tensor contraction, for gfx950 (CDNA4) MFMA.
//
// One kernel serves every level and impl:
//   * operands are gathered through device scatter vectors (the reference's block-scatter matrices,
//     src/matrix/stra_block_scatter_matrix.hpp) and the Strassen operand sums sum_i c_i A_i are
//     formed while staging the K-tile into LDS -- the fused packing of
//     src/stra_nodes/stra_packm.hpp:367-683 and the stra_pack_*_ukr of src/kernels/3m/gemm.hpp:505-2227;
//   * the register tile is accumulated with v_mfma_f64_16x16x4_f64 (or the f32 form for float);
//   * EPI_MULTI_C scatters the product into every C sub-block it feeds, C_i += alpha*coeff_i*M (the
//     multi-C micro-kernel of src/stra_nodes/stra_gemm_ukr.hpp:626-930), applying beta exactly once per
//     C sub-block; EPI_STORE writes the product to a staging buffer (AB and Naive impls).
// For ABC a workgroup owns one (tm,tn) tile position of every C sub-block and runs all 7^L products for
// it in order, so no two workgroups write the same C element and no atomics are needed.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "launch.hpp"

namespace mistra {

template <typename T>
struct Mfma;

template <>
struct Mfma<double>
{
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 run(double a, double b, v4 c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // output element v of a lane sits in row lane/16 + 4*v, column lane%16 (measured on gfx950)
    static __device__ __forceinline__ int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};

template <>
struct Mfma<float>
{
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 run(float a, float b, v4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // output element v of a lane sits in row 4*(lane/16) + v, column lane%16 (measured on gfx950)
    static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

// Tile index -> (tm, tn).  Workgroups are dispatched round-robin over the 8 XCDs (bid % 8), so the
// first remap gives each XCD a contiguous range of tiles; inside the range tiles are visited in
// groups of kGroupM tile-rows so the XCD's L2 holds a few A row-panels and B column-panels.
__device__ __forceinline__ void tile_coords(int bid, int tiles_m, int tiles_n, int& tm, int& tn)
{
    constexpr int kXcd = 8, kGroupM = 4;
    const int tiles = tiles_m * tiles_n;
    const int per = tiles / kXcd;
    int tile = bid < per * kXcd ? (bid % kXcd) * per + bid / kXcd : bid;
    const int in_group = kGroupM * tiles_n;
    const int group = tile / in_group;
    const int first_m = group * kGroupM;
    const int gsz = min(tiles_m - first_m, kGroupM);
    const int r = tile - group * in_group;
    tm = first_m + r % gsz;
    tn = r / gsz;
}

template <typename T, int TM, int TN, int TK, bool A_MFAST, bool B_NFAST, int EPI>
__global__ __launch_bounds__(256) void leaf_gemm_kernel(GemmArgs g)
{
    constexpr int NT = 256;
    constexpr int WM = TM / 2, WN = TN / 2;  // 2x2 waves
    constexpr int FM = WM / 16, FN = WN / 16;
    constexpr int LDA = TM + 1, LDB = TN + 1;  // +1 element: conflict-free K-fast LDS stores
    constexpr int EA = TM * TK / NT, EB = TN * TK / NT;
    static_assert(EA >= 1 && EB >= 1, "tile too small for 256 threads");
    static_assert(NT % TM == 0 && NT % TN == 0 && NT % TK == 0, "loader layout");
    typedef typename Mfma<T>::v4 v4;

    __shared__ T sA[2][TK * LDA];
    __shared__ T sB[2][TK * LDB];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm0 = (wave & 1) * WM;
    const int wn0 = (wave >> 1) * WN;

    int tm, tn;
    tile_coords(blockIdx.x, g.tiles_m, g.tiles_n, tm, tn);
    const int64_t row0 = int64_t(tm) * TM;
    const int64_t col0 = int64_t(tn) * TN;
    const int64_t ms = g.ms, ns = g.ns, ks = g.ks;

    const int64_t z = blockIdx.z;
    const T* __restrict__ Ag = static_cast<const T*>(g.A) + (g.batch_a ? g.batch_a[z] : 0);
    const T* __restrict__ Bg = static_cast<const T*>(g.B) + (g.batch_b ? g.batch_b[z] : 0);
    T* __restrict__ Cg = static_cast<T*>(g.C) + (g.batch_c ? g.batch_c[z] : 0);

    // Loader positions inside the tile.
    // A: M-fast -> consecutive lanes walk m (coalesced when m has unit stride); else lanes walk k.
    const int a_m = A_MFAST ? tid % TM : tid / TK;
    const int a_k = A_MFAST ? tid / TM : tid % TK;
    constexpr int A_MSTEP = A_MFAST ? 0 : NT / TK;
    constexpr int A_KSTEP = A_MFAST ? NT / TM : 0;
    const int b_n = B_NFAST ? tid % TN : tid / TK;
    const int b_k = B_NFAST ? tid / TN : tid % TK;
    constexpr int B_NSTEP = B_NFAST ? 0 : NT / TK;
    constexpr int B_KSTEP = B_NFAST ? NT / TN : 0;

    int pbeg = 0, pend = g.nprod;
    if (EPI == EPI_STORE)
    {
        pbeg = blockIdx.y;
        pend = pbeg + 1;
    }

    for (int p = pbeg; p < pend; p++)
    {
        const ProdDesc& P = g.prods[p];
        const int nA = P.nA, nB = P.nB;

        // Row offsets of A (per sub-block) and column offsets of B: fixed for the whole K loop.
        int64_t aro[kMaxTerms][A_MFAST ? 1 : EA];
        int64_t bco[kMaxTerms][B_NFAST ? 1 : EB];
        T acoef[kMaxTerms], bcoef[kMaxTerms];
#pragma unroll
        for (int a = 0; a < kMaxTerms; a++)
        {
            acoef[a] = a < nA ? T(P.a[a].coeff) : T(0);
#pragma unroll
            for (int e = 0; e < (A_MFAST ? 1 : EA); e++)
            {
                const int64_t r = row0 + a_m + e * A_MSTEP;
                aro[a][e] = (a < nA && r < ms) ? P.a[a].base + g.a_m[P.a[a].roff + r] : 0;
            }
        }
#pragma unroll
        for (int b = 0; b < kMaxTerms; b++)
        {
            bcoef[b] = b < nB ? T(P.b[b].coeff) : T(0);
#pragma unroll
            for (int e = 0; e < (B_NFAST ? 1 : EB); e++)
            {
                const int64_t c = col0 + b_n + e * B_NSTEP;
                bco[b][e] = (b < nB && c < ns) ? P.b[b].base + g.b_n[P.b[b].coff + c] : 0;
            }
        }

        T ra[kMaxTerms][EA], rb[kMaxTerms][EB];

        // Issue the global loads of one K-tile (raw values of every sub-block; summed later so the
        // loads stay in flight across the MFMA work of the previous tile).
        auto load_tile = [&](int64_t k0) {
#pragma unroll
            for (int a = 0; a < kMaxTerms; a++)
            {
                if (a < nA)
                {
                    const int64_t kb = P.a[a].coff + k0;
                    const int64_t run = g.a_krun[kb];
                    const int64_t c0 = g.a_k[kb];
#pragma unroll
                    for (int e = 0; e < EA; e++)
                    {
                        const int mm = a_m + e * A_MSTEP;
                        const int kk = a_k + e * A_KSTEP;
                        const bool ok = (row0 + mm < ms) && (k0 + kk < ks);
                        int64_t co = 0;
                        if (ok) co = run != 0 ? c0 + kk * run : g.a_k[kb + kk];
                        ra[a][e] = ok ? Ag[aro[a][A_MFAST ? 0 : e] + co] : T(0);
                    }
                }
            }
#pragma unroll
            for (int b = 0; b < kMaxTerms; b++)
            {
                if (b < nB)
                {
                    const int64_t kb = P.b[b].roff + k0;
                    const int64_t run = g.b_krun[kb];
                    const int64_t r0 = g.b_k[kb];
#pragma unroll
                    for (int e = 0; e < EB; e++)
                    {
                        const int nn = b_n + e * B_NSTEP;
                        const int kk = b_k + e * B_KSTEP;
                        const bool ok = (col0 + nn < ns) && (k0 + kk < ks);
                        int64_t ro = 0;
                        if (ok) ro = run != 0 ? r0 + kk * run : g.b_k[kb + kk];
                        rb[b][e] = ok ? Bg[bco[b][B_NFAST ? 0 : e] + ro] : T(0);
                    }
                }
            }
        };

        // Form the Strassen sums in registers and store the K-tile to LDS as [k][m] / [k][n].
        auto store_tile = [&](int buf) {
#pragma unroll
            for (int e = 0; e < EA; e++)
            {
                T v = acoef[0] * ra[0][e];
#pragma unroll
                for (int a = 1; a < kMaxTerms; a++)
                    if (a < nA) v += acoef[a] * ra[a][e];
                sA[buf][(a_k + e * A_KSTEP) * LDA + a_m + e * A_MSTEP] = v;
            }
#pragma unroll
            for (int e = 0; e < EB; e++)
            {
                T v = bcoef[0] * rb[0][e];
#pragma unroll
                for (int b = 1; b < kMaxTerms; b++)
                    if (b < nB) v += bcoef[b] * rb[b][e];
                sB[buf][(b_k + e * B_KSTEP) * LDB + b_n + e * B_NSTEP] = v;
            }
        };

        v4 acc[FM][FN];
#pragma unroll
        for (int i = 0; i < FM; i++)
#pragma unroll
            for (int j = 0; j < FN; j++) acc[i][j] = v4{0, 0, 0, 0};

        const int64_t nkt = (ks + TK - 1) / TK;
        if (nkt > 0)
        {
            load_tile(0);
            store_tile(0);
        }
        __syncthreads();

        for (int64_t kt = 0; kt < nkt; kt++)
        {
            const int cur = int(kt & 1);
            if (kt + 1 < nkt) load_tile((kt + 1) * TK);
            const T* __restrict__ la = sA[cur];
            const T* __restrict__ lb = sB[cur];
#pragma unroll
            for (int k4 = 0; k4 < TK / 4; k4++)
            {
                const int kr = k4 * 4 + (lane >> 4);
                T af[FM], bf[FN];
#pragma unroll
                for (int i = 0; i < FM; i++) af[i] = la[kr * LDA + wm0 + i * 16 + (lane & 15)];
#pragma unroll
                for (int j = 0; j < FN; j++) bf[j] = lb[kr * LDB + wn0 + j * 16 + (lane & 15)];
#pragma unroll
                for (int i = 0; i < FM; i++)
#pragma unroll
                    for (int j = 0; j < FN; j++) acc[i][j] = Mfma<T>::run(af[i], bf[j], acc[i][j]);
            }
            if (kt + 1 < nkt) store_tile(cur ^ 1);
            __syncthreads();
        }

        // Epilogue: element v of a lane's 16x16 block is (Mfma<T>::row(lane, v), lane % 16).
        if (EPI == EPI_MULTI_C)
        {
            const double alpha = g.alpha, beta = g.beta;
#pragma unroll
            for (int c = 0; c < kMaxTerms; c++)
            {
                if (c < P.nC)
                {
                    const SubRef cr = P.c[c];
                    const T coef = T(alpha * cr.coeff);
                    const bool first = (P.first_mask >> c) & 1;
                    const T bet = T(beta);
#pragma unroll
                    for (int j = 0; j < FN; j++)
                    {
                        const int64_t n = col0 + wn0 + j * 16 + (lane & 15);
                        if (n >= ns) continue;
                        const int64_t con = cr.base + g.c_n[cr.coff + n];
#pragma unroll
                        for (int i = 0; i < FM; i++)
                        {
#pragma unroll
                            for (int v = 0; v < 4; v++)
                            {
                                const int64_t m = row0 + wm0 + i * 16 + Mfma<T>::row(lane, v);
                                if (m >= ms) continue;
                                T* ptr = Cg + g.c_m[cr.roff + m] + con;
                                const T val = coef * acc[i][j][v];
                                if (first)
                                    *ptr = (beta == 0.0) ? val : bet * *ptr + val;
                                else
                                    *ptr += val;
                            }
                        }
                    }
                }
            }
        }
        else
        {
            T* Mp = static_cast<T*>(g.M) + (z * g.nprod + p) * ms * ns;
#pragma unroll
            for (int j = 0; j < FN; j++)
            {
                const int64_t n = col0 + wn0 + j * 16 + (lane & 15);
                if (n >= ns) continue;
#pragma unroll
                for (int i = 0; i < FM; i++)
#pragma unroll
                    for (int v = 0; v < 4; v++)
                    {
                        const int64_t m = row0 + wm0 + i * 16 + Mfma<T>::row(lane, v);
                        if (m < ms) Mp[n * ms + m] = acc[i][j][v];
                    }
            }
        }
    }
}

namespace {

template <typename T, int TM, int TN, int TK, int EPI>
hipError_t launch_tile(const GemmArgs& g, bool a_mfast, bool b_nfast, hipStream_t s, dim3 grid)
{
    if (a_mfast && b_nfast)
        hipLaunchKernelGGL((leaf_gemm_kernel<T, TM, TN, TK, true, true, EPI>), grid, dim3(256), 0, s, g);
    else if (a_mfast)
        hipLaunchKernelGGL((leaf_gemm_kernel<T, TM, TN, TK, true, false, EPI>), grid, dim3(256), 0, s, g);
    else if (b_nfast)
        hipLaunchKernelGGL((leaf_gemm_kernel<T, TM, TN, TK, false, true, EPI>), grid, dim3(256), 0, s, g);
    else
        hipLaunchKernelGGL((leaf_gemm_kernel<T, TM, TN, TK, false, false, EPI>), grid, dim3(256), 0, s, g);
    return hipGetLastError();
}

template <typename T, int EPI>
hipError_t launch_cfg(const GemmArgs& g, int tile, bool a_mfast, bool b_nfast, hipStream_t s, dim3 grid)
{
    constexpr int TK = sizeof(T) == 8 ? 16 : 32;
    if (tile == TILE_64)
        return launch_tile<T, 64, 64, TK, EPI>(g, a_mfast, b_nfast, s, grid);
    return launch_tile<T, 128, 128, TK, EPI>(g, a_mfast, b_nfast, s, grid);
}

}  // namespace

int tile_extent(int tile) { return tile == TILE_64 ? 64 : 128; }
int tile_kdepth(int tile, int elem_size)
{
    (void)tile;
    return elem_size == 8 ? 16 : 32;
}

hipError_t launch_leaf_gemm(int elem_size, int epi, int tile, GemmArgs g, bool a_mfast, bool b_nfast,
                            hipStream_t stream)
{
    const int te = tile_extent(tile);
    g.tiles_m = int((g.ms + te - 1) / te);
    g.tiles_n = int((g.ns + te - 1) / te);
    if (g.tiles_m == 0 || g.tiles_n == 0 || g.nbatch == 0) return hipSuccess;
    dim3 grid(unsigned(g.tiles_m) * unsigned(g.tiles_n), epi == EPI_STORE ? unsigned(g.nprod) : 1u,
              unsigned(g.nbatch));
    if (elem_size == 8)
        return epi == EPI_STORE ? launch_cfg<double, EPI_STORE>(g, tile, a_mfast, b_nfast, stream, grid)
                                : launch_cfg<double, EPI_MULTI_C>(g, tile, a_mfast, b_nfast, stream, grid);
    return epi == EPI_STORE ? launch_cfg<float, EPI_STORE>(g, tile, a_mfast, b_nfast, stream, grid)
                            : launch_cfg<float, EPI_MULTI_C>(g, tile, a_mfast, b_nfast, stream, grid);
}

}  // namespace mistra
