// Leaf-product GEMM kernel template of the Strassen tensor contraction, for gfx950 (CDNA4) MFMA.
// Instantiated per element type in leaf_f64.hip / leaf_f32.hip.
//
// One kernel serves every level and impl:
//   * operands are gathered through device scatter vectors (the reference's block-scatter matrices,
//     src/matrix/stra_block_scatter_matrix.hpp) and the Strassen operand sums sum_i c_i A_i are
//     formed while staging the K-tile into LDS -- the fused packing of
//     src/stra_nodes/stra_packm.hpp:367-683 and the stra_pack_*_ukr of src/kernels/3m/gemm.hpp:505-2227;
//   * the register tile is accumulated with v_mfma_f64_16x16x4_f64 (or the f32 form for float);
//   * EPI_MULTI_C scatters the product into every C sub-block it feeds, C_i += alpha*coeff_i*M (the
//     multi-C micro-kernel of src/stra_nodes/stra_gemm_ukr.hpp:626-930), applying beta exactly once per
//     C sub-block; EPI_STORE writes the product to a staging buffer (AB and Naive impls);
//     EPI_ATOMIC_C adds alpha*coeff_i*M into every target with FP64 atomic adds (beta applied by a pass
//     before the launch), so all products of a plan can share one launch.
// ABC otherwise issues the products as consecutive stage launches on one stream (the serialized
// products of src/stra_internal/3t/stra_mult.cxx:445-592), grouped so that no two workgroups of a
// launch write the same C element and no atomics are needed.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

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

// Kernel-start inputs (batch offsets) are read through the constant address space: with a uniform
// address that becomes a scalar load into SGPRs.  Inside the K loop scalar loads are avoided: they
// share the lgkmcnt counter with LDS reads and return out of order, so any wait for an LDS read would
// also wait for them.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(4))) T* cptr(const T* p)
{
    return (const __attribute__((address_space(4))) T*)(p);
}

// Operand and result pointers taken from tables in memory reach the kernel as generic (flat) pointers.
// Flat loads may target LDS, so the compiler must wait for both vmcnt and lgkmcnt to reach zero before
// using any of them, which leaves one load in flight per wave.  The data lives in device memory, so it
// is addressed through the global address space instead (in-order vmcnt: many loads in flight).
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gptr(T* p)
{
    return (__attribute__((address_space(1))) T*)(p);
}

// Device-coherent element access (agent scope, `sc1`): the load cannot hit a line of this XCD's L2 that
// another XCD has since written, and the store writes through this XCD's L2 to memory, so once it has
// completed (vmcnt) every XCD reads the new value.  The whole-L2 alternative, a release/acquire fence,
// writes back and invalidates the entire L2 of the XCD (buffer_wbl2 / buffer_inv), which stalls every
// workgroup on it.
template <typename P>
__device__ __forceinline__ auto coherent_load(P p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename P, typename T>
__device__ __forceinline__ void coherent_store(P p, T v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int64_t lmin(int64_t a, int64_t b) { return a < b ? a : b; }
constexpr int cmax(int a, int b) { return a > b ? a : b; }

// Multi-term products in the fast body: with one register set the loads of each term land during
// 1/TERMS of a tile's MFMA work; with two sets each load gets two phases.  Measured on gfx950 (128 x 128
// tile, FP64, 16-byte K loads): two sets are faster for 2 and 4 terms (at 2 terms 57.3 vs 54.5 TFLOP/s
// per leaf; before the 16-byte loads freed address registers, one set was faster).
template <int TERMS, int KMODE>
constexpr bool two_sets()
{
    // 2 terms: only where the 16-byte loads leave the registers for it (other modes would spill)
    return TERMS >= 4 || (TERMS == 2 && KMODE == KMODE_AFFINE_VEC);
}

// Raw buffer loads for the fast body: a uniform base (the SGPR resource) plus a 32-bit lane offset,
// one instruction per element.  An offset at or beyond kOob reads 0 (the resource's range check),
// which zero-fills rows/columns past the tile edge and K past ks without branches or selects.
constexpr uint32_t kOob = 0x80000000u;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* base)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, int(kOob), 0x00020000);
}
// kOob when k > klim, else 0 (OR-ed into an in-range offset, which is below kOob)
__device__ __forceinline__ uint32_t past(int k, int klim) { return uint32_t(klim - k) & kOob; }
template <typename T>
__device__ __forceinline__ T bload(__amdgpu_buffer_rsrc_t r, uint32_t off);
// 16 bytes (W elements of T) in one load
template <typename T, int W>
__device__ __forceinline__ void bload16(__amdgpu_buffer_rsrc_t r, uint32_t off, T* out)
{
    static_assert(W * sizeof(T) == 16, "16-byte load");
    auto x = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
    __builtin_memcpy(out, &x, 16);
}
template <>
__device__ __forceinline__ double bload<double>(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
    auto x = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
    double d;
    __builtin_memcpy(&d, &x, sizeof(d));
    return d;
}
template <>
__device__ __forceinline__ float bload<float>(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
    auto x = __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0);
    float f;
    __builtin_memcpy(&f, &x, sizeof(f));
    return f;
}

// Lane `l`'s copy of a 64-bit value, in a scalar register.
__device__ __forceinline__ int64_t readlane64(int64_t x, int l)
{
    const int lo = __builtin_amdgcn_readlane(int(uint64_t(x) & 0xffffffffu), l);
    const int hi = __builtin_amdgcn_readlane(int(uint64_t(x) >> 32), l);
    return int64_t((uint64_t(uint32_t(hi)) << 32) | uint64_t(uint32_t(lo)));
}

// A value the same in every lane, held as such (in SGPRs): a no-op for values the compiler already knows
// to be uniform (kernel arguments and scalar loads); for ones it loaded with vector loads (arguments
// read from a table after a fence) it keeps buffer resources built from them out of waterfall loops.
__device__ __forceinline__ int64_t uniform64(int64_t x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const int lo = __builtin_amdgcn_readfirstlane(int(uint64_t(x) & 0xffffffffu));
    const int hi = __builtin_amdgcn_readfirstlane(int(uint64_t(x) >> 32));
    return int64_t((uint64_t(uint32_t(hi)) << 32) | uint64_t(uint32_t(lo)));
#else
    return x;
#endif
}

// The body of one workgroup: register tile bx of the launch's tile grid, product by, batch entry bz.
// at_tm >= 0: the tile at (at_tm, at_tn) of the grid instead (callers that choose the tile themselves).
template <typename T, int TM, int TN, int TK, int TERMS, bool A_MFAST, bool B_NFAST, int EPI, int KMODE>
__device__ __forceinline__ void leaf_gemm_body(const GemmArgs& g, int bx, int by, int64_t bz, int at_tm = -1,
                                               int at_tn = 0)
{
    constexpr int NT = 256;
    constexpr int WM = TM / 2, WN = TN / 2;  // 2x2 waves
    constexpr int FM = WM / 16, FN = WN / 16;
    constexpr int LDA = TM + 1, LDB = TN + 1;  // +1 element: conflict-free K-fast LDS stores
    constexpr int EA = TM * TK / NT, EB = TN * TK / NT;  // elements per thread and K-tile
    // 16-byte loads along unit-stride K (KMODE_AFFINE_VEC, K-fast operands): WA (WB) consecutive k per load
    constexpr bool VA = KMODE == KMODE_AFFINE_VEC && !A_MFAST, VB = KMODE == KMODE_AFFINE_VEC && !B_NFAST;
    constexpr int WA = VA ? 16 / int(sizeof(T)) : 1, WB = VB ? 16 / int(sizeof(T)) : 1;
    constexpr int LA = EA / WA, LB = EB / WB;                      // loads per thread and K-tile
    constexpr int RA = A_MFAST ? 1 : LA, RB = B_NFAST ? 1 : LB;    // row (col) positions per thread
    static_assert(EA >= 1 && EB >= 1, "tile too small for 256 threads");
    static_assert(NT % TM == 0 && NT % TN == 0 && NT % TK == 0, "loader layout");
    typedef typename Mfma<T>::v4 v4;

    // Direct-to-LDS K-tiles (fast body, see below): single-term launches, both operands K-fast, 128-byte
    // K rows.
    constexpr int ROWB = TK * int(sizeof(T));
    constexpr bool kDma = KMODE == KMODE_AFFINE_VEC && !A_MFAST && !B_NFAST && TERMS == 1 && ROWB == 128 &&
                          (TM * ROWB) % (NT * 16) == 0 && (TN * ROWB) % (NT * 16) == 0;
    // K buffers; the multi-C epilogue reuses the A buffers as a 32-row staging tile (32 x (TN + 1)
    // elements), or the B buffers when only those hold it (TN > TM)
    constexpr int SAN0 = cmax(TK * LDA, kDma ? TERMS * TM * TK : 0);
    constexpr int SBN = cmax(TK * LDB, kDma ? TERMS * TN * TK : 0);
    constexpr bool kStageB = 16 * (TN + 1) > SAN0 && 16 * (TN + 1) <= SBN;
    constexpr int SAN = kStageB ? SAN0 : cmax(SAN0, 16 * (TN + 1));
    __shared__ __attribute__((aligned(16))) T sA[2][SAN];
    __shared__ __attribute__((aligned(16))) T sB[2][SBN];
    constexpr bool kMultiC = EPI == EPI_MULTI_C || EPI == EPI_ATOMIC_C;
    __shared__ int64_t ccol[kMultiC ? kMaxTerms : 1][TN];  // C column offsets per target

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    // the same in every lane of a wave: made a scalar, so the per-wave LDS addresses of the direct-to-LDS
    // loads (M0) and the fragment offsets are scalar arithmetic rather than vector adds and read-backs
#if defined(__HIP_DEVICE_COMPILE__)
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
#else
    const int wave = tid >> 6;
#endif
    const int wm0 = (wave & 1) * WM;
    const int wn0 = (wave >> 1) * WN;

    int tm = at_tm, tn = at_tn;
    if (at_tm < 0) tile_coords(bx, g.tiles_m, g.tiles_n, tm, tn);
    const int64_t row0 = int64_t(tm) * TM;
    const int64_t col0 = int64_t(tn) * TN;
    const int64_t ms = g.ms, ns = g.ns, ks = g.ks;

    const int64_t z = bz;
    const T* __restrict__ Ag = static_cast<const T*>(g.A) + (g.batch_a ? cptr(g.batch_a)[z] : 0);
    const T* __restrict__ Bg = static_cast<const T*>(g.B) + (g.batch_b ? cptr(g.batch_b)[z] : 0);
    __attribute__((address_space(1))) T* __restrict__ Cg = gptr(static_cast<T*>(g.C)) + (g.batch_c ? cptr(g.batch_c)[z] : 0);
    const T* __restrict__ zero = static_cast<const T*>(g.zero);

    // Loader positions inside the tile.  A: M-fast -> consecutive lanes walk m (coalesced when m has
    // unit stride); else lanes walk k.  B likewise with n.
    const int a_m = A_MFAST ? tid % TM : tid / (TK / WA);
    const int a_k = A_MFAST ? tid / TM : (tid % (TK / WA)) * WA;
    constexpr int A_MSTEP = A_MFAST ? 0 : NT * WA / TK;
    constexpr int A_KSTEP = A_MFAST ? NT / TM : 0;
    const int b_n = B_NFAST ? tid % TN : tid / (TK / WB);
    const int b_k = B_NFAST ? tid / TN : (tid % (TK / WB)) * WB;
    constexpr int B_NSTEP = B_NFAST ? 0 : NT * WB / TK;
    constexpr int B_KSTEP = B_NFAST ? NT / TN : 0;

    // Epilogue, shared by both bodies.  Element v of a lane's 16x16 accumulator block is
    // (Mfma<T>::row(lane, v), lane % 16).  `translate`: C targets are shifts (ProdDesc base) of the first
    // C sub-block, else they are addressed through (roff, coff) into the scatter vectors.
    // `ab`: a factor (+-1) the operands were not scaled by (the direct-to-LDS loads load them as stored)
    auto epilogue = [&](const ProdDesc& P, int p, auto& acc, bool translate, double ab) {
        if constexpr (kMultiC)
        {
            // Multi-C update through LDS: for each 16-row block i, the waves write their accumulators
            // (straight from the accumulator registers) into a 32 x TN staging tile, then all 256
            // threads update every C target from it, consecutive threads on consecutive rows (C's
            // unit-stride dimension after the plan's orientation).  Keeps the accumulators out of the
            // VGPRs, so the main loop keeps a small register budget.
            constexpr int LDE = TN + 1;
            constexpr int RROWS = 32, CSTEP = NT / RROWS, CPER = TN / CSTEP;
            static_assert(2 * 16 * LDE <= 2 * (kStageB ? SBN : SAN), "staging tile must fit its buffers");
            T* stage = kStageB ? sB[0] : sA[0];  // the K-tile buffers are free after the K loop's last barrier
            const double alpha = g.alpha, beta = g.beta;
            const int nC = P.nC;
            const int first_mask = P.first_mask;
            const bool coherent = g.c_coherent != 0;
            // column offsets of every target, for this tile (one shared row in translate mode)
            for (int x = tid; x < (translate ? 1 : nC) * TN; x += NT)
            {
                const int c = x / TN, col = x % TN;
                ccol[c][col] = g.c_n[(translate ? 0 : P.c[c].coff) + lmin(col0 + col, ns - 1)];
            }
            const int rl = tid % RROWS, cg = tid / RROWS;
#pragma unroll 1
            for (int i = 0; i < FM; i++)
            {
                // static accumulator indices (a uniform branch per block) keep acc in registers
#pragma unroll
                for (int ii = 0; ii < FM; ii++)
                    if (ii == i)
                    {
#pragma unroll
                        for (int j = 0; j < FN; j++)
#pragma unroll
                            for (int v = 0; v < 4; v++)
                                stage[((wave & 1) * 16 + Mfma<T>::row(lane, v)) * LDE + wn0 + j * 16 + (lane & 15)] =
                                    acc[ii][j][v];
                    }
                __syncthreads();
                const int64_t grow = row0 + (rl / 16) * WM + i * 16 + (rl % 16);
                const bool rok = grow < ms;
                const int64_t growc = lmin(grow, ms - 1);
                const int64_t rbase = translate ? g.c_m[growc] : 0;
#pragma unroll 1
                for (int c = 0; c < nC; c++)
                {
                    const T coef = T(alpha * P.c[c].coeff * ab);
                    if constexpr (EPI == EPI_ATOMIC_C)
                    {
                        // C already scaled by beta: one hardware atomic add per element (the products of the
                        // launch that share this target add in an unspecified order)
                        T* Crow = (T*)(Cg + P.c[c].base + (translate ? rbase : g.c_m[P.c[c].roff + growc]));
                        const int64_t* cc = ccol[translate ? 0 : c];
#pragma unroll
                        for (int e = 0; e < CPER; e++)
                        {
                            const int col = cg + e * CSTEP;
                            if (rok && col0 + col < ns) unsafeAtomicAdd(&Crow[cc[col]], coef * stage[rl * LDE + col]);
                        }
                        continue;
                    }
                    const bool first = (first_mask >> c) & 1;
                    const bool read_c = !(first && beta == 0.0);
                    const T bet = first ? T(beta) : T(1);
                    __attribute__((address_space(1))) T* Crow = Cg + P.c[c].base + (translate ? rbase : g.c_m[P.c[c].roff + growc]);
                    const int64_t* cc = ccol[translate ? 0 : c];
                    T old[CPER];
                    if (coherent)
                    {
#pragma unroll
                        for (int e = 0; e < CPER; e++) old[e] = read_c ? coherent_load(&Crow[cc[cg + e * CSTEP]]) : T(0);
                    }
                    else
                    {
#pragma unroll
                        for (int e = 0; e < CPER; e++) old[e] = read_c ? Crow[cc[cg + e * CSTEP]] : T(0);
                    }
#pragma unroll
                    for (int e = 0; e < CPER; e++)
                    {
                        const int col = cg + e * CSTEP;
                        const T val = coef * stage[rl * LDE + col];
                        const T out = read_c ? bet * old[e] + val : val;
                        if (rok && col0 + col < ns)
                        {
                            if (coherent)
                                coherent_store(&Crow[cc[col]], out);
                            else
                                Crow[cc[col]] = out;
                        }
                    }
                }
                __syncthreads();
            }
        }
        else
        {
            (void)P;
            (void)translate;
            const T sab = T(ab);
            T* Mp = static_cast<T*>(g.M) + (z * g.nprod + p) * ms * ns;
#pragma unroll
            for (int i = 0; i < FM; i++)
#pragma unroll
                for (int v = 0; v < 4; v++)
                {
                    const int64_t m = row0 + wm0 + i * 16 + Mfma<T>::row(lane, v);
#pragma unroll
                    for (int j = 0; j < FN; j++)
                    {
                        const int64_t n = col0 + wn0 + j * 16 + (lane & 15);
                        if (m < ms && n < ns) Mp[n * ms + m] = sab * acc[i][j][v];
                    }
                }
        }
    };

    if constexpr (KMODE != KMODE_GENERAL)
    {
        // ---- fast body -------------------------------------------------------------------------------
        // Chosen by the host (kfast_check) when every sum term's sub-block is a translate of the first
        // sub-block in all three scatter vectors of its operand: term t of product p then sits at a
        // uniform shift (ProdDesc base, precomputed on the host) from the first sub-block.
        //   KMODE_AFFINE: the K runs of the first sub-block are affine at every K-tile start with stride
        //     g.ka_stride (A) / g.kb_stride (B), so an element's byte offset within a K-tile is the same
        //     for all terms and K-tiles: a load is uniform base (shift + the tile's first K offset) +
        //     that offset.
        //   KMODE_GATHER: any K layout: each lane reads the K offsets of its own K positions for the
        //     next tile (a_k / b_k, one small load each), a load is uniform base (shift) + row offset +
        //     K offset.
        constexpr bool KG = KMODE == KMODE_GATHER;
        const int p = by;
        const ProdDesc& P = g.prods[p];
        const int nA = __builtin_amdgcn_readfirstlane(P.nA), nB = __builtin_amdgcn_readfirstlane(P.nB);

        // Per-lane byte offsets of the K-tile elements, relative to the tile's first K offset; rows
        // (columns) past the tile edge get an out-of-range offset and read 0.
        uint32_t aoff[LA], boff[LB];
        {
            int64_t ar[RA], bc[RB];
            bool aok[RA], bok[RB];
#pragma unroll
            for (int e = 0; e < RA; e++)
            {
                aok[e] = row0 + a_m + e * A_MSTEP < ms;
                ar[e] = g.a_m[lmin(row0 + a_m + e * A_MSTEP, ms - 1)];
            }
#pragma unroll
            for (int e = 0; e < RB; e++)
            {
                bok[e] = col0 + b_n + e * B_NSTEP < ns;
                bc[e] = g.b_n[lmin(col0 + b_n + e * B_NSTEP, ns - 1)];
            }
#pragma unroll
            for (int e = 0; e < LA; e++)
                aoff[e] = aok[A_MFAST ? 0 : e]
                              ? uint32_t((ar[A_MFAST ? 0 : e] + (KG ? 0 : int64_t(a_k + e * A_KSTEP) * g.ka_stride)) *
                                         int64_t(sizeof(T)))
                              : kOob;
#pragma unroll
            for (int e = 0; e < LB; e++)
                boff[e] = bok[B_NFAST ? 0 : e]
                              ? uint32_t((bc[B_NFAST ? 0 : e] + (KG ? 0 : int64_t(b_k + e * B_KSTEP) * g.kb_stride)) *
                                         int64_t(sizeof(T)))
                              : kOob;
        }
        T acoef[TERMS], bcoef[TERMS];
        int64_t ash[TERMS], bsh[TERMS];
#pragma unroll
        for (int t = 0; t < TERMS; t++)
        {
            acoef[t] = t < nA ? T(P.a[t].coeff) : T(0);
            bcoef[t] = t < nB ? T(P.b[t].coeff) : T(0);
            ash[t] = uniform64(t < nA ? P.a[t].base : 0);
            bsh[t] = uniform64(t < nB ? P.b[t].base : 0);
        }

        const int64_t nkt = (ks + TK - 1) / TK;
        // K-tile counters are 32-bit (the host rejects K of 2^31 tiles or more): the scalar unit has no
        // signed 64-bit compare, so 64-bit tile counters put the loop tests and clamps on the vector
        // unit, which does not run beside the MFMAs
        const int nkt32 = int(nkt);
        // K offsets of a tile.  Affine: the tile's first K offsets (lane 0: A, lane 1: B), one vector
        // load, read into SGPRs with v_readlane.  Gathered: each lane's K offsets for its own K
        // positions (positions past ks clamp to ks - 1; their elements read 0).  Fetched one tile
        // before the tile's loads are issued.
        constexpr int NKA = KG ? (A_MFAST ? EA : 1) : 1, NKB = KG ? (B_NFAST ? EB : 1) : 1;
        typedef std::conditional_t<KG, uint32_t, int64_t> KT;  // gathered offsets fit 31 bits (host-checked)
        struct KOff
        {
            KT a[NKA], b[NKB];
        };
        int64_t kinfo = 0;
        KOff kgat;
        auto fetch_kinfo = [&](int kt) {
            const int kc = kt < nkt32 - 1 ? kt : nkt32 - 1;
            if constexpr (KG)
            {
                // 32-bit positions and byte offsets (the host limits gathered K to ks < 2^28): a load from a
                // scalar base plus a lane offset, with no 64-bit vector arithmetic
                const int k0 = kc * TK, kmax = int(ks) - 1;
                const char* ak = reinterpret_cast<const char*>(g.a_k);
                const char* bk = reinterpret_cast<const char*>(g.b_k);
#pragma unroll
                for (int i = 0; i < NKA; i++)
                {
                    const int kk = min(k0 + a_k + i * A_KSTEP, kmax);
                    kgat.a[i] = KT(*reinterpret_cast<const int64_t*>(ak + uint32_t(kk) * 8u));
                }
#pragma unroll
                for (int i = 0; i < NKB; i++)
                {
                    const int kk = min(k0 + b_k + i * B_KSTEP, kmax);
                    kgat.b[i] = KT(*reinterpret_cast<const int64_t*>(bk + uint32_t(kk) * 8u));
                }
            }
            else
                kinfo = (lane == 1 ? g.b_k : g.a_k)[int64_t(kc) * TK];
        };
        auto read_kinfo = [&]() {
            if constexpr (KG)
                return kgat;
            else
            {
                KOff k;
                k.a[0] = readlane64(kinfo, 0);
                k.b[0] = readlane64(kinfo, 1);
                return k;
            }
        };

        // Load jobs: job (tile kt, term t) reads term t of K-tile kt into one of two register sets;
        // the K loop runs TERMS phases per tile, and phase q issues job q + TERMS + 1 and adds job
        // q + TERMS (issued one phase earlier) into the next tile's LDS buffer -- term 0 stores, later
        // terms read-modify-write their own elements, so phases need no barrier.  Every load has two
        // phases of MFMA work to land in, and the register cost is two single-term sets whatever
        // the number of terms.
        constexpr bool kTwoSets = two_sets<TERMS, KMODE>();
        T ra[kTwoSets ? 2 : 1][EA], rb[kTwoSets ? 2 : 1][EB];
        // Every issue executes its loads (a job past the last tile or a term the product lacks reads
        // out of range, i.e. zeros with no memory traffic): the load count between any two points of
        // the loop is then the same on every path, so each wait covers exactly the loads it needs
        // (a skipped load on one path would make the waits count conservatively, down to vmcnt(0)).
        auto issue = [&](auto set, int t, int kt, const KOff& ko) {
            constexpr int S = decltype(set)::value;
            const int klim = kt < nkt32 ? int(lmin(ks - 1 - int64_t(kt) * TK, TK - 1)) : -1;  // k > klim reads 0
            const int kla = t < nA ? klim : -1, klb = t < nB ? klim : -1;
            const __amdgpu_buffer_rsrc_t ra_ = make_rsrc(Ag + ash[t] + (KG ? 0 : ko.a[0]));
#pragma unroll
            for (int e = 0; e < LA; e++)
            {
                const uint32_t kof = KG ? uint32_t(ko.a[A_MFAST ? e : 0]) * uint32_t(sizeof(T)) : 0u;  // < 2^31
                const uint32_t off = (aoff[e] + kof) | past(a_k + e * A_KSTEP, kla);
                if constexpr (VA)
                    bload16<T, WA>(ra_, off, &ra[S][e * WA]);  // ks is a multiple of TK: whole vectors
                else
                    ra[S][e] = bload<T>(ra_, off);
            }
            const __amdgpu_buffer_rsrc_t rb_ = make_rsrc(Bg + bsh[t] + (KG ? 0 : ko.b[0]));
#pragma unroll
            for (int e = 0; e < LB; e++)
            {
                const uint32_t kof = KG ? uint32_t(ko.b[B_NFAST ? e : 0]) * uint32_t(sizeof(T)) : 0u;
                const uint32_t off = (boff[e] + kof) | past(b_k + e * B_KSTEP, klb);
                if constexpr (VB)
                    bload16<T, WB>(rb_, off, &rb[S][e * WB]);
                else
                    rb[S][e] = bload<T>(rb_, off);
            }
        };
        // Term sums: term 0 stores its scaled values into the LDS buffer, later terms read-modify-write
        // their own elements (a running sum in registers measured slower: it costs the registers that
        // hide load latency).
        // A single term is stored as loaded and its coefficient (+-1) applied in the epilogue, as on the
        // direct-to-LDS path: no multiply per element in the loop (vector instructions there take MFMA
        // issue time)
        auto put = [&](auto set, int t, int buf) {
            constexpr int S = decltype(set)::value;
            if (t < nA)
            {
#pragma unroll
                for (int e = 0; e < LA; e++)
#pragma unroll
                    for (int j = 0; j < WA; j++)
                    {
                        T& dst = sA[buf][(a_k + j + e * A_KSTEP) * LDA + a_m + e * A_MSTEP];
                        const T v = TERMS == 1 ? ra[S][e * WA + j] : acoef[t] * ra[S][e * WA + j];
                        dst = t == 0 ? v : dst + v;
                    }
            }
            if (t < nB)
            {
#pragma unroll
                for (int e = 0; e < LB; e++)
#pragma unroll
                    for (int j = 0; j < WB; j++)
                    {
                        T& dst = sB[buf][(b_k + j + e * B_KSTEP) * LDB + b_n + e * B_NSTEP];
                        const T v = TERMS == 1 ? rb[S][e * WB + j] : bcoef[t] * rb[S][e * WB + j];
                        dst = t == 0 ? v : dst + v;
                    }
            }
        };
        typedef std::integral_constant<int, 0> S0;
        typedef std::integral_constant<int, kTwoSets ? 1 : 0> S1;

        v4 acc[FM][FN];
#pragma unroll
        for (int i = 0; i < FM; i++)
#pragma unroll
            for (int j = 0; j < FN; j++) acc[i][j] = v4{0, 0, 0, 0};

        constexpr int K4 = TK / 4;
        static_assert(TERMS <= K4 && K4 % TERMS == 0, "one phase per term");
        constexpr int KPH = K4 / TERMS;  // k4 steps per phase
        auto mfma_steps = [&](const T* __restrict__ la, const T* __restrict__ lb, int k4_begin, int k4_end) {
#pragma unroll
            for (int k4 = k4_begin; k4 < k4_end; k4++)
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
        };

        // Single-term launches with both operands K-fast: K-tiles go from memory straight into LDS
        // (buffer_load ... lds), with no register staging and no LDS store instructions.  A wave's load
        // writes 64 lanes x 16 bytes contiguously, so the LDS tile is [row][k] (one 128-byte row per M
        // or N position) with the 16-byte chunks of row r XOR-swizzled by r & 7: the MFMA fragment
        // reads (16 consecutive rows at one k) spread each 8 consecutive rows over the 8 chunks of a
        // 128-byte bank window.  The operands are loaded as stored, so the term coefficients (+-1) move
        // to the epilogue.
        //
        // The loop keeps the buffer index a run-time value, and the fragment reads that follow a load
        // into LDS in program order sit in another basic block than the load: that way the compiler
        // leaves the reads without a wait for the loads, while an unrolled loop with constant buffers,
        // reads merged into ds_read2st64, or reads in the loads' own block all made it wait, or issue
        // the loads late, and cost 8-13%.  Correctness never depends on this: an explicit vmcnt(0)
        // before each barrier that publishes a tile waits for the loads.
        if constexpr (kDma)
        {
            typedef __attribute__((address_space(3))) void* lds_ptr;
            constexpr int CH = 16 / int(sizeof(T));                                     // elements per chunk
            constexpr int NIA = TM * ROWB / (NT * 16), NIB = TN * ROWB / (NT * 16);   // loads per thread
            static_assert(TM * TK <= SAN && TN * TK <= SBN, "DMA tiles fit the K buffers");
            const int lr = lane >> 3;        // row within the wave's 8-row block
            const int gch = (lane & 7) ^ lr; // chunk of the row this lane's LDS slot holds
            uint32_t adma[NIA], bdma[NIB];
#pragma unroll
            for (int i = 0; i < NIA; i++)
            {
                const int64_t r = row0 + (i * 4 + wave) * 8 + lr;
                adma[i] = r < ms ? uint32_t((g.a_m[r] + gch * CH) * int64_t(sizeof(T))) : kOob;
            }
#pragma unroll
            for (int i = 0; i < NIB; i++)
            {
                const int64_t c = col0 + (i * 4 + wave) * 8 + lr;
                bdma[i] = c < ns ? uint32_t((g.b_n[c] + gch * CH) * int64_t(sizeof(T))) : kOob;
            }
            // ks is a multiple of TK (KMODE_AFFINE_VEC): every K-tile is whole
            auto dma = [&](int buf, const KOff& ko) {
#if defined(__HIP_DEVICE_COMPILE__)  // the builtin exists for the gfx950 target only, not the host pass
                const __amdgpu_buffer_rsrc_t ra_ = make_rsrc(Ag + ash[0] + ko.a[0]);
                const __amdgpu_buffer_rsrc_t rb_ = make_rsrc(Bg + bsh[0] + ko.b[0]);
#pragma unroll
                for (int i = 0; i < NIA; i++)
                    __builtin_amdgcn_raw_ptr_buffer_load_lds(
                        ra_, (lds_ptr)(reinterpret_cast<char*>(sA[buf]) + (i * 4 + wave) * 1024), 16, adma[i], 0, 0, 0);
#pragma unroll
                for (int i = 0; i < NIB; i++)
                    __builtin_amdgcn_raw_ptr_buffer_load_lds(
                        rb_, (lds_ptr)(reinterpret_cast<char*>(sB[buf]) + (i * 4 + wave) * 1024), 16, bdma[i], 0, 0, 0);
#else
                (void)buf;
                (void)ko;
#endif
            };
            // Every barrier that publishes loaded tiles is preceded by this wait for the wave's own loads
            // into LDS (vmcnt counts them).  It is written out rather than left to the compiler, whose
            // wait insertion does not reliably see a later LDS read depending on an earlier load into LDS
            // across loop iterations.
            auto dma_landed = [&]() {
#if defined(__HIP_DEVICE_COMPILE__)
                __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0); expcnt and lgkmcnt unconstrained
#endif
            };
            // The MFMA sums over whichever k each lane group supplies, as long as A and B agree: lane
            // group g (lane / 16) takes chunk 4 grp + g of the row for the CH consecutive MFMA steps of
            // group grp, element e of the chunk in step CH grp + e, so one 16-byte LDS read feeds CH
            // MFMA steps (the 16 k of a K-tile are still each used once).
            typedef T vec16 __attribute__((ext_vector_type(CH)));
            auto mfma_dma = [&](const T* __restrict__ la, const T* __restrict__ lb) {
                const int sw = lane & 7;  // row & 7 of every fragment row this lane reads
                static_assert(K4 % CH == 0 && (K4 / CH) * 4 * CH == TK, "chunks of a K-tile");
#pragma unroll
                for (int grp = 0; grp < K4 / CH; grp++)
                {
                    const int ko = ((grp * 4 + (lane >> 4)) ^ sw) * CH;
                    vec16 af[FM], bf[FN];
#pragma unroll
                    for (int i = 0; i < FM; i++)
                        af[i] = *reinterpret_cast<const vec16*>(&la[(wm0 + i * 16 + (lane & 15)) * TK + ko]);
#pragma unroll
                    for (int j = 0; j < FN; j++)
                        bf[j] = *reinterpret_cast<const vec16*>(&lb[(wn0 + j * 16 + (lane & 15)) * TK + ko]);
#pragma unroll
                    for (int e = 0; e < CH; e++)
#pragma unroll
                        for (int i = 0; i < FM; i++)
#pragma unroll
                            for (int j = 0; j < FN; j++) acc[i][j] = Mfma<T>::run(af[i][e], bf[j][e], acc[i][j]);
                }
            };
            if constexpr (K4 / CH == 2)
            {
                // Two fragment groups per K-tile, software-pipelined.  Iteration kt reads group 1 of
                // tile kt while group 0's MFMAs run; after the barrier (tile kt fully read, tile kt + 1
                // landed) it reads group 0 of tile kt + 1 and loads tile kt + 2 into the buffer tile kt
                // leaves, while group 1's MFMAs run.  A load thus has a whole iteration (64 MFMAs per
                // wave) to land before the barrier that publishes it, and no MFMA waits for an LDS read
                // it could have had earlier.  The loads sit in their own block, after the reads: a read
                // after a load into LDS in the same block makes the compiler wait for the load (it cannot
                // tell the buffers apart).  After the last barrier the read of the other buffer is stale
                // and unused (it stays unconditional: no branch around reads).
                auto read_grp = [&](const T* __restrict__ la, const T* __restrict__ lb, int grp, vec16* af,
                                    vec16* bf) {
                    const int ko = ((grp * 4 + (lane >> 4)) ^ (lane & 7)) * CH;
#pragma unroll
                    for (int i = 0; i < FM; i++)
                        af[i] = *reinterpret_cast<const vec16*>(&la[(wm0 + i * 16 + (lane & 15)) * TK + ko]);
#pragma unroll
                    for (int j = 0; j < FN; j++)
                        bf[j] = *reinterpret_cast<const vec16*>(&lb[(wn0 + j * 16 + (lane & 15)) * TK + ko]);
                };
                auto mma_grp = [&](const vec16* af, const vec16* bf) {
#pragma unroll
                    for (int e = 0; e < CH; e++)
#pragma unroll
                        for (int i = 0; i < FM; i++)
#pragma unroll
                            for (int j = 0; j < FN; j++) acc[i][j] = Mfma<T>::run(af[i][e], bf[j][e], acc[i][j]);
                };
                fetch_kinfo(0);
                dma(0, read_kinfo());
                fetch_kinfo(1);
                dma(1, read_kinfo());  // with one tile the offsets clamp to it: a repeat, waited for below
                fetch_kinfo(2);
                dma_landed();
                __syncthreads();
                vec16 a0[FM], b0[FN], a1[FM], b1[FN];
                read_grp(sA[0], sB[0], 0, a0, b0);
                for (int kt = 0; kt < nkt32; kt++)
                {
                    const int cur = kt & 1;
                    read_grp(sA[cur], sB[cur], 1, a1, b1);
                    mma_grp(a0, b0);
                    // group 1's reads go between group 0's first MFMAs (each MFMA occupies the pipe for
                    // many cycles; the wave issues the reads in its shadow)
#pragma unroll
                    for (int i = 0; i < FM + FN; i++)
                    {
                        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
                        __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // DS read
                    }
                    // scheduling fences keep the rest in this order: left alone, the scheduler sinks
                    // the reads after the barrier into the next iteration and hoists the barrier (and
                    // its wait for all reads) above group 0's MFMAs
                    __builtin_amdgcn_sched_barrier(0);
                    dma_landed();
                    __syncthreads();
                    read_grp(sA[cur ^ 1], sB[cur ^ 1], 0, a0, b0);
                    __builtin_amdgcn_sched_barrier(0);
                    const KOff kn = read_kinfo();
                    fetch_kinfo(kt + 3);
                    if (kt + 2 < nkt32) dma(cur, kn);
                    mma_grp(a1, b1);
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
            else
            {
                fetch_kinfo(0);
                dma(0, read_kinfo());
                fetch_kinfo(1);
                dma_landed();
                __syncthreads();
                for (int kt = 0; kt < nkt32; kt++)
                {
                    const int cur = kt & 1;
                    const KOff kn = read_kinfo();
                    fetch_kinfo(kt + 2);
                    // buffer cur ^ 1 was last read in iteration kt - 1, before its closing barrier
                    if (kt + 1 < nkt32) dma(cur ^ 1, kn);
                    mfma_dma(sA[cur], sB[cur]);
                    dma_landed();
                    __syncthreads();
                }
            }
            epilogue(P, p, acc, g.c_scatter == 0, double(acoef[0]) * double(bcoef[0]));
            return;
        }
        else if constexpr (!kTwoSets)
        {
            // One register set: phase t of tile kt issues term t of tile kt + 1, runs 1/TERMS of the
            // tile's MFMA work and adds the term into the next LDS buffer.
            fetch_kinfo(0);
            {
                const KOff k0 = read_kinfo();
#pragma unroll
                for (int t = 0; t < TERMS; t++)
                {
                    issue(S0(), t, 0, k0);
                    put(S0(), t, 0);
                }
            }
            fetch_kinfo(1);
            __syncthreads();
            for (int kt = 0; kt < nkt32; kt++)
            {
                const int cur = kt & 1;
                const KOff kn = read_kinfo();
                fetch_kinfo(kt + 2);
#pragma unroll
                for (int ph = 0; ph < TERMS; ph++)
                {
                    issue(S0(), ph, kt + 1, kn);
                    mfma_steps(sA[cur], sB[cur], ph * KPH, (ph + 1) * KPH);
                    if (kt + 1 < nkt32) put(S0(), ph, cur ^ 1);
                }
                __syncthreads();
            }
        }
        else
        {
            // prologue: tile 0 straight into buffer 0, job (1, 0) in flight, tile 2's offsets requested
            fetch_kinfo(0);
            KOff kcur = read_kinfo();
#pragma unroll
            for (int t = 0; t < TERMS; t++)
            {
                issue(S0(), t, 0, kcur);
                put(S0(), t, 0);
            }
            fetch_kinfo(1);
            kcur = read_kinfo();
            fetch_kinfo(2);  // before the job's loads, as in the loop, so the first wait for it is exact
            issue(S0(), 0, 1, kcur);  // job TERMS: set TERMS & 1 = 0
            __syncthreads();
            static_assert(TERMS % 2 == 0, "register-set choice must not depend on the tile parity");
            for (int kt = 0; kt < nkt32; kt++)
            {
                const int cur = kt & 1;
                const KOff knext = read_kinfo();
                fetch_kinfo(kt + 3);
#pragma unroll
                for (int ph = 0; ph < TERMS; ph++)
                {
                    // phase q = kt * TERMS + ph issues job q + TERMS + 1 into set (ph + 1) & 1 --
                    // (kt + 1, ph + 1), or (kt + 2, 0) in the last phase -- and puts job q + TERMS =
                    // (kt + 1, ph) from set ph & 1
                    if (ph + 1 < TERMS)
                    {
                        if ((ph & 1) == 0) issue(S1(), ph + 1, kt + 1, kcur);
                        else issue(S0(), ph + 1, kt + 1, kcur);
                    }
                    else
                        issue(S0(), 0, kt + 2, knext);
                    mfma_steps(sA[cur], sB[cur], ph * KPH, (ph + 1) * KPH);
                    if (kt + 1 < nkt32)
                    {
                        if ((ph & 1) == 0) put(S0(), ph, cur ^ 1);
                        else put(S1(), ph, cur ^ 1);
                    }
                }
                kcur = knext;
                __syncthreads();
            }
        }
        epilogue(P, p, acc, true, TERMS == 1 ? double(acoef[0]) * double(bcoef[0]) : 1.0);
        return;
    }
    else
    {
    // One product per workgroup: `by` selects it among the launch's products.  (ABC launches
        // one product at a time, in stream order, so that products updating the same C sub-block are
        // serialized; a loop over products inside the kernel costs registers in the main loop.)
        {
            const int p = by;
            const ProdDesc& P = g.prods[p];
            const int nA = __builtin_amdgcn_readfirstlane(P.nA), nB = __builtin_amdgcn_readfirstlane(P.nB);
    
            // Row offsets of A (per sub-block) and column offsets of B: fixed for the whole K loop.
            // Positions outside the leaf are clamped to a valid one and flagged; such lanes later read
            // the zero block, so every load is issued unconditionally and its value used as is.
            int64_t aro[TERMS][RA], bco[TERMS][RB];
            bool arow_ok[RA], bcol_ok[RB];
            T acoef[TERMS], bcoef[TERMS];
    #pragma unroll
            for (int e = 0; e < RA; e++) arow_ok[e] = row0 + a_m + e * A_MSTEP < ms;
    #pragma unroll
            for (int e = 0; e < RB; e++) bcol_ok[e] = col0 + b_n + e * B_NSTEP < ns;
    #pragma unroll
            for (int a = 0; a < TERMS; a++)
            {
                const bool on = a < nA;
                acoef[a] = on ? T(P.a[a].coeff) : T(0);
                const int64_t base = on ? P.a[a].base : 0, roff = on ? P.a[a].roff : 0;
    #pragma unroll
                for (int e = 0; e < RA; e++)
                    aro[a][e] = base + g.a_m[roff + lmin(row0 + a_m + e * A_MSTEP, ms - 1)];
            }
    #pragma unroll
            for (int b = 0; b < TERMS; b++)
            {
                const bool on = b < nB;
                bcoef[b] = on ? T(P.b[b].coeff) : T(0);
                const int64_t base = on ? P.b[b].base : 0, coff = on ? P.b[b].coff : 0;
    #pragma unroll
                for (int e = 0; e < RB; e++)
                    bco[b][e] = base + g.b_n[coff + lmin(col0 + b_n + e * B_NSTEP, ns - 1)];
            }
    
            // K-run info of a K-tile.  Lane t < TERMS holds the block-scatter stride of A term t's run
            // starting at the tile (0 = not affine), lane TERMS+t the offset of its first element, lanes
            // 2*TERMS+t and 3*TERMS+t the same for B.  One vector load per tile fetches all of it, one tile
            // ahead of use (vector loads retire in order, so waiting for a tile's data never waits for
            // it), and v_readlane moves each entry to a scalar register when it is used.  (Loads from a
            // uniform address would become scalar loads, whose out-of-order lgkmcnt would make every LDS
            // wait of the MFMA block wait for them too.)
            const int kw = lane < 4 * TERMS ? lane / TERMS : 3;
            const int kt_term = lane % TERMS;
            const int64_t kpos = kw < 2 ? (kt_term < nA ? P.a[kt_term].coff : 0) : (kt_term < nB ? P.b[kt_term].roff : 0);
            const int64_t* ksrc = kw == 0 ? g.a_krun : kw == 1 ? g.a_k : kw == 2 ? g.b_krun : g.b_k;
            int64_t kinfo = 0;
            auto fetch_kinfo = [&](int64_t k0) { kinfo = ksrc[kpos + k0]; };
    
            T ra[TERMS][EA], rb[TERMS][EB];
    
            // Issue the global loads of one K-tile (raw values of every sub-block; they are summed after
            // the MFMA work of the previous tile, so the loads stay in flight across it).  The affine and
            // gathered forms are whole uniform branches, so no per-element control flow (and no wait at
            // its join) is generated.
            auto load_tile = [&](int64_t k0) {
                const int klim = int(lmin(ks - 1 - k0, TK - 1));  // last valid k inside this tile
                // All run info first (its load is a tile old, so nothing newer is waited for), then one
                // uniform branch: every term affine (the common case) or the general per-term code.
                int64_t arun[TERMS], ac0[TERMS], brun[TERMS], br0[TERMS];
                bool affine = true;
    #pragma unroll
                for (int a = 0; a < TERMS; a++)
                {
                    arun[a] = readlane64(kinfo, a);
                    ac0[a] = readlane64(kinfo, TERMS + a);
                    affine = affine && (a >= nA || arun[a] != 0);
                }
    #pragma unroll
                for (int b = 0; b < TERMS; b++)
                {
                    brun[b] = readlane64(kinfo, 2 * TERMS + b);
                    br0[b] = readlane64(kinfo, 3 * TERMS + b);
                    affine = affine && (b >= nB || brun[b] != 0);
                }
                if (affine)
                {
    #pragma unroll
                    for (int a = 0; a < TERMS; a++)
                    {
                        if (a < nA)
                        {
    #pragma unroll
                            for (int e = 0; e < EA; e++)
                            {
                                const int kk = a_k + e * A_KSTEP;
                                const int kc = kk < klim ? kk : klim;
                                const bool ok = arow_ok[A_MFAST ? 0 : e] && kk <= klim;
                                ra[a][e] = *(ok ? Ag + aro[a][A_MFAST ? 0 : e] + ac0[a] + int64_t(kc) * arun[a] : zero);
                            }
                        }
                    }
    #pragma unroll
                    for (int b = 0; b < TERMS; b++)
                    {
                        if (b < nB)
                        {
    #pragma unroll
                            for (int e = 0; e < EB; e++)
                            {
                                const int kk = b_k + e * B_KSTEP;
                                const int kc = kk < klim ? kk : klim;
                                const bool ok = bcol_ok[B_NFAST ? 0 : e] && kk <= klim;
                                rb[b][e] = *(ok ? Bg + bco[b][B_NFAST ? 0 : e] + br0[b] + int64_t(kc) * brun[b] : zero);
                            }
                        }
                    }
                    return;
                }
    #pragma unroll
                for (int a = 0; a < TERMS; a++)
                {
                    if (a < nA)
                    {
                        const int64_t run = arun[a];
                        if (run != 0)
                        {
                            const int64_t c0 = ac0[a];
    #pragma unroll
                            for (int e = 0; e < EA; e++)
                            {
                                const int kk = a_k + e * A_KSTEP;
                                const int kc = kk < klim ? kk : klim;
                                const bool ok = arow_ok[A_MFAST ? 0 : e] && kk <= klim;
                                ra[a][e] = *(ok ? Ag + aro[a][A_MFAST ? 0 : e] + c0 + int64_t(kc) * run : zero);
                            }
                        }
                        else
                        {
                            const int64_t kb = readlane64(kpos, TERMS + a) + k0;
                            int64_t co[EA];
    #pragma unroll
                            for (int e = 0; e < EA; e++)
                            {
                                const int kk = a_k + e * A_KSTEP;
                                co[e] = g.a_k[kb + (kk < klim ? kk : klim)];
                            }
    #pragma unroll
                            for (int e = 0; e < EA; e++)
                            {
                                const bool ok = arow_ok[A_MFAST ? 0 : e] && a_k + e * A_KSTEP <= klim;
                                ra[a][e] = *(ok ? Ag + aro[a][A_MFAST ? 0 : e] + co[e] : zero);
                            }
                        }
                    }
                }
    #pragma unroll
                for (int b = 0; b < TERMS; b++)
                {
                    if (b < nB)
                    {
                        const int64_t run = brun[b];
                        if (run != 0)
                        {
                            const int64_t r0 = br0[b];
    #pragma unroll
                            for (int e = 0; e < EB; e++)
                            {
                                const int kk = b_k + e * B_KSTEP;
                                const int kc = kk < klim ? kk : klim;
                                const bool ok = bcol_ok[B_NFAST ? 0 : e] && kk <= klim;
                                rb[b][e] = *(ok ? Bg + bco[b][B_NFAST ? 0 : e] + r0 + int64_t(kc) * run : zero);
                            }
                        }
                        else
                        {
                            const int64_t kb = readlane64(kpos, 3 * TERMS + b) + k0;
                            int64_t ro[EB];
    #pragma unroll
                            for (int e = 0; e < EB; e++)
                            {
                                const int kk = b_k + e * B_KSTEP;
                                ro[e] = g.b_k[kb + (kk < klim ? kk : klim)];
                            }
    #pragma unroll
                            for (int e = 0; e < EB; e++)
                            {
                                const bool ok = bcol_ok[B_NFAST ? 0 : e] && b_k + e * B_KSTEP <= klim;
                                rb[b][e] = *(ok ? Bg + bco[b][B_NFAST ? 0 : e] + ro[e] : zero);
                            }
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
                    for (int a = 1; a < TERMS; a++)
                        if (a < nA) v += acoef[a] * ra[a][e];
                    sA[buf][(a_k + e * A_KSTEP) * LDA + a_m + e * A_MSTEP] = arow_ok[A_MFAST ? 0 : e] ? v : T(0);
                }
    #pragma unroll
                for (int e = 0; e < EB; e++)
                {
                    T v = bcoef[0] * rb[0][e];
    #pragma unroll
                    for (int b = 1; b < TERMS; b++)
                        if (b < nB) v += bcoef[b] * rb[b][e];
                    sB[buf][(b_k + e * B_KSTEP) * LDB + b_n + e * B_NSTEP] = bcol_ok[B_NFAST ? 0 : e] ? v : T(0);
                }
            };
    
            v4 acc[FM][FN];
    #pragma unroll
            for (int i = 0; i < FM; i++)
    #pragma unroll
                for (int j = 0; j < FN; j++) acc[i][j] = v4{0, 0, 0, 0};
    
            const int64_t nkt = (ks + TK - 1) / TK;
            fetch_kinfo(0);
            load_tile(0);
            if (nkt > 1) fetch_kinfo(TK);
            store_tile(0);
            __syncthreads();
    
            for (int64_t kt = 0; kt < nkt; kt++)
            {
                const int cur = int(kt & 1);
                if (kt + 1 < nkt)
                {
                    load_tile((kt + 1) * TK);
                    if (kt + 2 < nkt) fetch_kinfo((kt + 2) * TK);
                }
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
    
            epilogue(P, p, acc, false, 1.0);
        }
    }
}

// One launch: grid (tiles, products, batch entries).
template <typename T, int TM, int TN, int TK, int TERMS, bool A_MFAST, bool B_NFAST, int EPI, int KMODE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void leaf_gemm_kernel(GemmArgs g)
{
    leaf_gemm_body<T, TM, TN, TK, TERMS, A_MFAST, B_NFAST, EPI, KMODE>(g, blockIdx.x, blockIdx.y, blockIdx.z);
}

// Grouped launch (batched contractions): job j owns workgroups [first[j], first[j + 1]) of a 1-D grid,
// laid out as (tile, product, batch entry) of its own GemmArgs; jobs of one launch write disjoint C.
template <typename T, int TM, int TN, int TK, int TERMS, bool A_MFAST, bool B_NFAST, int EPI, int KMODE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void leaf_gemm_grouped(
    const GemmArgs* __restrict__ jobs, const int64_t* __restrict__ first, int njobs)
{
    const int64_t b = blockIdx.x;
    int lo = 0, hi = njobs - 1;
    while (lo < hi)  // last job with first[j] <= b
    {
        const int mid = (lo + hi + 1) >> 1;
        if (first[mid] <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    const GemmArgs& g = jobs[lo];
    const int64_t local = b - first[lo];
    const int64_t tiles = int64_t(g.tiles_m) * g.tiles_n;
    const int64_t rest = local / tiles;
    leaf_gemm_body<T, TM, TN, TK, TERMS, A_MFAST, B_NFAST, EPI, KMODE>(g, int(local % tiles), int(rest % g.nprod),
                                                                      rest / g.nprod);
}

}  // namespace mistra
