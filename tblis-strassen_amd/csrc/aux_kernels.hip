// HBM-bound helper kernels of the Strassen tensor contraction (gfx950):
//   * operand sums S_p = sum_i c_i A_i for the Naive impl (the gather-sum `add` of
//     src/stra_internal/3t/tensor2matrix.hpp:821-825 used by straprim_naive2,
//     src/stra_internal/3t/straprim_common.hpp:70-157);
//   * the combine C_i = beta*C_i + alpha*sum_p coeff_{p,i} M_p for the AB and Naive impls (the
//     scatter-add of src/stra_internal/3t/tensor2matrix.hpp:827-830, here done once for all products);
//   * synthetic data, a calibration copy, a relative-error reduction and C = beta*C.
#include <hip/hip_runtime.h>

#include "launch.hpp"

namespace mistra {

// ---- operand sums ---------------------------------------------------------------------------------

template <typename T, bool ROWS_FAST>
__global__ __launch_bounds__(256) void operand_sums_kernel(SumArgs a)
{
    constexpr int TR = 64, TC = 64;
    __shared__ T tile[TC][TR + 1];
    const int tiles_r = int((a.rows + TR - 1) / TR);
    const int tr = blockIdx.x % tiles_r, tc = blockIdx.x / tiles_r;
    const int p = blockIdx.y;
    const ProdDesc& P = a.prods[p];
    const int nt = a.which == 0 ? P.nA : P.nB;
    const SubRef* terms = a.which == 0 ? P.a : P.b;
    const T* X = static_cast<const T*>(a.X) + a.zoff;
    const int tid = threadIdx.x;
    const int64_t r0 = int64_t(tr) * TR, c0 = int64_t(tc) * TC;

    // read phase: lanes along the unit-stride dimension of the source
#pragma unroll
    for (int e = 0; e < (TR * TC) / 256; e++)
    {
        const int lin = tid + e * 256;
        const int rr = ROWS_FAST ? lin % TR : lin / TC;
        const int cc = ROWS_FAST ? lin / TR : lin % TC;
        const int64_t r = r0 + rr, c = c0 + cc;
        T v = T(0);
        if (r < a.rows && c < a.cols)
        {
#pragma unroll
            for (int t = 0; t < kMaxTerms; t++)
                if (t < nt)
                    v += T(terms[t].coeff) * X[terms[t].base + a.rs[terms[t].roff + r] + a.cs[terms[t].coff + c]];
        }
        tile[cc][rr] = v;
    }
    __syncthreads();
    // write phase: dense column-major, rows fastest
    T* out = static_cast<T*>(a.out) + int64_t(p) * a.rows * a.cols;
#pragma unroll
    for (int e = 0; e < (TR * TC) / 256; e++)
    {
        const int lin = tid + e * 256;
        const int rr = lin % TR, cc = lin / TR;
        const int64_t r = r0 + rr, c = c0 + cc;
        if (r < a.rows && c < a.cols) out[c * a.rows + r] = tile[cc][rr];
    }
}

// Translate-form sums along unit-stride rows: each thread owns RPT consecutive rows (one row-vector
// load for the whole tile) and TC / CG columns; a term costs one RPT-wide load per column.
template <typename T, bool PAIR>
__global__ __launch_bounds__(256) void operand_sums_rows_kernel(SumArgs a)
{
    constexpr int RPT = PAIR ? 2 : 1;          // rows per thread
    constexpr int TR = 256;                    // rows per tile
    constexpr int RT = TR / RPT;               // threads along the rows
    constexpr int CG = 256 / RT;               // column groups
    constexpr int TC = 16;                     // columns per tile
    typedef T vec2 __attribute__((ext_vector_type(2)));
    const int64_t tiles_r = (a.rows + TR - 1) / TR;
    const int64_t tr = blockIdx.x % tiles_r, tc = blockIdx.x / tiles_r;
    const int p = blockIdx.y;
    const ProdDesc& P = a.prods[p];
    const int nt = a.which == 0 ? P.nA : P.nB;
    const SubRef* terms = a.which == 0 ? P.a : P.b;
    const int tid = threadIdx.x;
    const int64_t r = tr * TR + int64_t(tid % RT) * RPT;
    if (r >= a.rows) return;
    const T* X = static_cast<const T*>(a.X) + a.zoff + a.rs[r];
    T* out = static_cast<T*>(a.out) + int64_t(p) * a.rows * a.cols + r;
    int64_t base[kMaxTerms];
    T coeff[kMaxTerms];
#pragma unroll
    for (int t = 0; t < kMaxTerms; t++)
    {
        base[t] = t < nt ? terms[t].base : 0;
        coeff[t] = t < nt ? T(terms[t].coeff) : T(0);
    }
#pragma unroll
    for (int j = 0; j < TC / CG; j++)
    {
        const int64_t c = tc * TC + tid / RT + int64_t(j) * CG;
        if (c >= a.cols) break;
        const int64_t cb = a.cs[c];
        if constexpr (PAIR)
        {
            vec2 acc = {T(0), T(0)};
#pragma unroll
            for (int t = 0; t < kMaxTerms; t++)
                if (t < nt)
                {
                    const vec2 v = *reinterpret_cast<const vec2*>(X + base[t] + cb);
                    acc.x += coeff[t] * v.x;
                    acc.y += coeff[t] * v.y;
                }
            *reinterpret_cast<vec2*>(out + c * a.rows) = acc;
        }
        else
        {
            T acc = T(0);
#pragma unroll
            for (int t = 0; t < kMaxTerms; t++)
                if (t < nt) acc += coeff[t] * X[base[t] + cb];
            out[c * a.rows] = acc;
        }
    }
}

// One pass over the sources for all staged sums (MultiSumArgs): a thread owns RPT consecutive rows of
// one column, loads its NSRC source values once and writes every output from registers.
template <typename T, bool PAIR, int NSRC, bool BYVAL>
__global__ __launch_bounds__(256) void operand_multisums_kernel(MultiSumArgs a)
{
    constexpr int RPT = PAIR ? 2 : 1;
    typedef T vec2 __attribute__((ext_vector_type(2)));
    const int64_t tiles_r = (a.rows + 256 * RPT - 1) / (256 * RPT);
    const int64_t tr = blockIdx.x % tiles_r, c = blockIdx.x / tiles_r;
    const int64_t r = (tr * 256 + threadIdx.x) * RPT;
    if (r >= a.rows) return;
    const T* X = static_cast<const T*>(a.X) + a.zoff + a.rs[r] + a.cs[c];
    T v[NSRC][RPT];
#pragma unroll
    for (int s = 0; s < NSRC; s++)
    {
        const T* x = X + a.shift[s];
        if constexpr (PAIR)
        {
            const vec2 w = *reinterpret_cast<const vec2*>(x);
            v[s][0] = w.x;
            v[s][1] = w.y;
        }
        else
            v[s][0] = x[0];
    }
    T* out = static_cast<T*>(a.out) + c * a.rows + r;
    const int64_t stride = a.rows * a.cols;
    for (int q = 0; q < a.nout; q++)
    {
        T acc[RPT];
#pragma unroll
        for (int i = 0; i < RPT; i++) acc[i] = T(0);
#pragma unroll
        for (int s = 0; s < NSRC; s++)
        {
            const double cf = BYVAL ? a.cv[q * NSRC + s] : a.coef[q * NSRC + s];  // uniform
            if (cf != 0.0)
#pragma unroll
                for (int i = 0; i < RPT; i++) acc[i] += T(cf) * v[s][i];
        }
        if constexpr (PAIR)
            *reinterpret_cast<vec2*>(out + q * stride) = vec2{acc[0], acc[1]};
        else
            out[q * stride] = acc[0];
    }
}

// The 2-level one-pass sums through the Kronecker factorization of their coefficients (MultiSumArgs::
// coef1): source (a, b) -- outer sub-block a, inner b -- is sub-block 4 (2 (a / 2) + b / 2) + 2 (a % 2) +
// b % 2 of the 4 x 4 split.
template <typename T, bool PAIR, bool BYVAL>
__global__ __launch_bounds__(256) void operand_multisums_kron_kernel(MultiSumArgs a)
{
    constexpr int RPT = PAIR ? 2 : 1;
    typedef T vec2 __attribute__((ext_vector_type(2)));
    const int64_t tiles_r = (a.rows + 256 * RPT - 1) / (256 * RPT);
    const int64_t tr = blockIdx.x % tiles_r, c = blockIdx.x / tiles_r;
    const int64_t r = (tr * 256 + threadIdx.x) * RPT;
    if (r >= a.rows) return;
    const T* X = static_cast<const T*>(a.X) + a.zoff + a.rs[r] + a.cs[c];
    T v[16][RPT];
#pragma unroll
    for (int s = 0; s < 16; s++)
    {
        const T* x = X + a.shift[s];
        if constexpr (PAIR)
        {
            const vec2 w = *reinterpret_cast<const vec2*>(x);
            v[s][0] = w.x;
            v[s][1] = w.y;
        }
        else
            v[s][0] = x[0];
    }
    T* out = static_cast<T*>(a.out) + c * a.rows + r;
    const int64_t stride = a.rows * a.cols;
    for (int o = 0; o < 7; o++)
    {
        T u[4][RPT];
#pragma unroll
        for (int b = 0; b < 4; b++)
        {
#pragma unroll
            for (int i = 0; i < RPT; i++) u[b][i] = T(0);
#pragma unroll
            for (int q = 0; q < 4; q++)  // outer sub-block
            {
                const double cf = BYVAL ? a.cv[o * 4 + q] : a.coef1[o * 4 + q];  // uniform
                if (cf != 0.0)
                {
                    const int src = 4 * (2 * (q / 2) + b / 2) + 2 * (q % 2) + b % 2;
#pragma unroll
                    for (int i = 0; i < RPT; i++) u[b][i] += T(cf) * v[src][i];
                }
            }
        }
        for (int p = 0; p < 7; p++)
        {
            T acc[RPT];
#pragma unroll
            for (int i = 0; i < RPT; i++) acc[i] = T(0);
#pragma unroll
            for (int b = 0; b < 4; b++)
            {
                const double cf = BYVAL ? a.cv[p * 4 + b] : a.coef1[p * 4 + b];
                if (cf != 0.0)
#pragma unroll
                    for (int i = 0; i < RPT; i++) acc[i] += T(cf) * u[b][i];
            }
            T* o_ = out + (7 * o + p) * stride;
            if constexpr (PAIR)
                *reinterpret_cast<vec2*>(o_) = vec2{acc[0], acc[1]};
            else
                o_[0] = acc[0];
        }
    }
}

template <typename T, bool PAIR, bool BYVAL>
hipError_t multisums_launch(const MultiSumArgs& a, dim3 grid, hipStream_t s)
{
    if (a.nsrc == 1) hipLaunchKernelGGL((operand_multisums_kernel<T, PAIR, 1, BYVAL>), grid, dim3(256), 0, s, a);
    else if (a.nsrc == 4) hipLaunchKernelGGL((operand_multisums_kernel<T, PAIR, 4, BYVAL>), grid, dim3(256), 0, s, a);
    else if (a.nsrc == 16 && a.coef1 && a.nout == 49)
        hipLaunchKernelGGL((operand_multisums_kron_kernel<T, PAIR, BYVAL>), grid, dim3(256), 0, s, a);
    else if (a.nsrc == 16) hipLaunchKernelGGL((operand_multisums_kernel<T, PAIR, 16, BYVAL>), grid, dim3(256), 0, s, a);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

template <typename T>
hipError_t multisums_typed(const MultiSumArgs& a, hipStream_t s)
{
    const int rpt = a.pair ? 2 : 1;
    const int64_t blocks = ((a.rows + 256 * rpt - 1) / (256 * rpt)) * a.cols;
    if (blocks == 0 || a.nout == 0) return hipSuccess;
    const dim3 grid(static_cast<unsigned>(blocks));
    const bool kron = a.nsrc == 16 && a.coef1 && a.nout == 49;
    const bool byval = kron ? a.ncv == 28 : (a.ncv > 0 && a.ncv == a.nsrc * a.nout);
    if (a.pair) return byval ? multisums_launch<T, true, true>(a, grid, s) : multisums_launch<T, true, false>(a, grid, s);
    return byval ? multisums_launch<T, false, true>(a, grid, s) : multisums_launch<T, false, false>(a, grid, s);
}

hipError_t launch_operand_multisums(int elem_size, const MultiSumArgs& a, hipStream_t s)
{
    return elem_size == 8 ? multisums_typed<double>(a, s) : multisums_typed<float>(a, s);
}

hipError_t launch_operand_sums(int elem_size, const SumArgs& a, hipStream_t s)
{
    if (a.shifted && a.rows_fast)
    {
        const int64_t tiles = ((a.rows + 255) / 256) * ((a.cols + 15) / 16);
        if (tiles == 0 || a.nprod == 0) return hipSuccess;
        dim3 grid(unsigned(tiles), unsigned(a.nprod));
        if (elem_size == 8)
        {
            if (a.pair) hipLaunchKernelGGL((operand_sums_rows_kernel<double, true>), grid, dim3(256), 0, s, a);
            else hipLaunchKernelGGL((operand_sums_rows_kernel<double, false>), grid, dim3(256), 0, s, a);
        }
        else
        {
            if (a.pair) hipLaunchKernelGGL((operand_sums_rows_kernel<float, true>), grid, dim3(256), 0, s, a);
            else hipLaunchKernelGGL((operand_sums_rows_kernel<float, false>), grid, dim3(256), 0, s, a);
        }
        return hipGetLastError();
    }
    const int64_t tiles = ((a.rows + 63) / 64) * ((a.cols + 63) / 64);
    if (tiles == 0 || a.nprod == 0) return hipSuccess;
    dim3 grid(unsigned(tiles), unsigned(a.nprod));
    if (elem_size == 8)
    {
        if (a.rows_fast) hipLaunchKernelGGL((operand_sums_kernel<double, true>), grid, dim3(256), 0, s, a);
        else hipLaunchKernelGGL((operand_sums_kernel<double, false>), grid, dim3(256), 0, s, a);
    }
    else
    {
        if (a.rows_fast) hipLaunchKernelGGL((operand_sums_kernel<float, true>), grid, dim3(256), 0, s, a);
        else hipLaunchKernelGGL((operand_sums_kernel<float, false>), grid, dim3(256), 0, s, a);
    }
    return hipGetLastError();
}

// ---- combine ----------------------------------------------------------------------------------------

// One workgroup covers kCombineRows x (4 * kCombineCols) elements of one C sub-block: each wave takes 64
// consecutive rows (M is row-fast, so every load of a wave is one 512-byte run for FP64) and each lane
// kCombineCols columns, whose loads are issued together.  No integer division per element; the row's C
// offset is read once per lane, the column offsets once per column.  The terms of a sub-block are summed
// in table order in FP64, the same arithmetic as the per-element form it replaces.
constexpr int kCombineRows = 64;
constexpr int kCombineCols = 8;

// Offset of element (r, c) of a product, and the product stride, in M's layout (CombineArgs::m_tiled);
// columns c .. c + j of one tile (c % kCombineCols == 0, j < kCombineCols) are kMTileRows apart.
template <bool TILED>
__device__ __forceinline__ int64_t m_offset(const CombineArgs& a, int64_t r, int64_t c)
{
    if constexpr (TILED)
        return ((c / kMTileCols) * (a.m_mp / kMTileRows) + r / kMTileRows) * (kMTileRows * kMTileCols) +
               (c % kMTileCols) * kMTileRows + r % kMTileRows;
    else
        return r + c * a.ms;
}
template <bool TILED>
__device__ __forceinline__ int64_t m_colstep(const CombineArgs& a)
{
    return TILED ? int64_t(kMTileRows) : a.ms;
}
template <bool TILED>
__device__ __forceinline__ int64_t m_slab(const CombineArgs& a)
{
    return TILED ? a.m_mp * a.m_np : a.ms * a.ns;
}

template <typename T, typename TM, bool TILED>
__global__ __launch_bounds__(256) void combine_kernel(CombineArgs a)
{
    const int s = blockIdx.y;
    const int64_t nel = m_slab<TILED>(a), cstep = m_colstep<TILED>(a);
    const int64_t row_tiles = (a.ms + kCombineRows - 1) / kCombineRows;
    const int64_t rt = int64_t(blockIdx.x) % row_tiles, ct = int64_t(blockIdx.x) / row_tiles;
    const int64_t r = rt * kCombineRows + (threadIdx.x & (kCombineRows - 1));
    const int64_t c0 = (ct * 4 + (threadIdx.x >> 6)) * kCombineCols;
    if (r >= a.ms || c0 >= a.ns) return;  // no barrier below
    const int ncol = a.ns - c0 < kCombineCols ? int(a.ns - c0) : kCombineCols;
    const int tb = a.term_begin[s], te = a.term_begin[s + 1];
    const int64_t roff = a.sub_roff[s], coff = a.sub_coff[s];
    const int64_t cm = a.c_m[roff + r];
    for (int64_t z = blockIdx.z; z < a.nbatch; z += gridDim.z)
    {
        T* C = static_cast<T*>(a.C) + (a.batch_c ? a.batch_c[z] : 0) + cm;
        const TM* M = static_cast<const TM*>(a.M) + z * int64_t(a.nprod) * nel + m_offset<TILED>(a, r, c0);
        double acc[kCombineCols];
#pragma unroll
        for (int j = 0; j < kCombineCols; j++) acc[j] = 0.0;
        if (ncol == kCombineCols)
        {
            for (int t = tb; t < te; t++)
            {
                const double w = a.terms[t].coeff;
                const TM* Mp = M + int64_t(a.terms[t].prod) * nel;
                TM v[kCombineCols];
#pragma unroll
                for (int j = 0; j < kCombineCols; j++) v[j] = Mp[int64_t(j) * cstep];
#pragma unroll
                for (int j = 0; j < kCombineCols; j++) acc[j] += w * double(v[j]);
            }
        }
        else
        {
            for (int t = tb; t < te; t++)
            {
                const double w = a.terms[t].coeff;
                const TM* Mp = M + int64_t(a.terms[t].prod) * nel;
#pragma unroll
                for (int j = 0; j < kCombineCols; j++)
                    if (j < ncol) acc[j] += w * double(Mp[int64_t(j) * cstep]);
            }
        }
#pragma unroll
        for (int j = 0; j < kCombineCols; j++)
        {
            if (j >= ncol) break;
            T* ptr = C + a.c_n[coff + c0 + j];
            const double v = a.alpha * acc[j];
            *ptr = a.beta == 0.0 ? T(v) : T(a.beta * double(*ptr) + v);
        }
    }
}

// All sub-blocks in one pass (CombineArgs::pmask, NSUB = nsub): a lane reads M_p at its NC positions once
// per product and adds it to every sub-block the product feeds, so M is read once rather than once per C
// target (12 of 7 products' worth at one level, 144 of 49 at two).  Products are visited in order and a
// sub-block the product does not feed is skipped (a uniform branch on the kernel-argument mask), so each
// sub-block sums the same terms in the same order as the per-sub-block form.  The loads of kCombineChunk
// products are issued together before any is used (a chunk past the last product re-reads the last one:
// in bounds, never added), so a wave keeps kCombineChunk * NC loads in flight.  With one product in flight
// the 2-level combine ran at 3.9 TB/s, latency-bound.
constexpr int kCombineChunk = 7;
template <typename T, int NSUB, int NC, typename TM, bool TILED>
__global__ __launch_bounds__(256) void combine_all_kernel(CombineArgs a)
{
    const int64_t nel = m_slab<TILED>(a);
    const int64_t row_tiles = (a.ms + kCombineRows - 1) / kCombineRows;
    const int64_t rt = int64_t(blockIdx.x) % row_tiles, ct = int64_t(blockIdx.x) / row_tiles;
    const int64_t r = rt * kCombineRows + (threadIdx.x & (kCombineRows - 1));
    const int64_t c0 = (ct * 4 + (threadIdx.x >> 6)) * NC;
    if (r >= a.ms || c0 >= a.ns) return;  // no barrier below
    const int ncol = a.ns - c0 < NC ? int(a.ns - c0) : NC;
    const int last = a.nprod - 1;
    for (int64_t z = blockIdx.z; z < a.nbatch; z += gridDim.z)
    {
        const TM* M = static_cast<const TM*>(a.M) + z * int64_t(a.nprod) * nel + m_offset<TILED>(a, r, c0);
        double acc[NSUB][NC];
#pragma unroll
        for (int s = 0; s < NSUB; s++)
#pragma unroll
            for (int j = 0; j < NC; j++) acc[s][j] = 0.0;
        // columns past the edge read column c0 again (in bounds) and are never stored
        int64_t coff_j[NC];
#pragma unroll
        for (int j = 0; j < NC; j++) coff_j[j] = (j < ncol ? j : 0) * m_colstep<TILED>(a);
        for (int p0 = 0; p0 < a.nprod; p0 += kCombineChunk)
        {
            TM v[kCombineChunk][NC];
#pragma unroll
            for (int u = 0; u < kCombineChunk; u++)
            {
                const TM* Mp = M + int64_t(p0 + u < last ? p0 + u : last) * nel;
#pragma unroll
                for (int j = 0; j < NC; j++) v[u][j] = Mp[coff_j[j]];
            }
#pragma unroll
            for (int u = 0; u < kCombineChunk; u++)
            {
                if (p0 + u > last) break;
                const unsigned pm = a.pmask[p0 + u];
#pragma unroll
                for (int s = 0; s < NSUB; s++)
                {
                    if (pm & (1u << s))
                    {
                        const double ws = (pm & (1u << (16 + s))) ? -1.0 : 1.0;
#pragma unroll
                        for (int j = 0; j < NC; j++) acc[s][j] += ws * double(v[u][j]);
                    }
                }
            }
        }
        T* C = static_cast<T*>(a.C) + (a.batch_c ? a.batch_c[z] : 0);
#pragma unroll
        for (int s = 0; s < NSUB; s++)
        {
            T* Cs = C + a.c_m[a.sub_roff[s] + r];
            const int64_t coff = a.sub_coff[s] + c0;
#pragma unroll
            for (int j = 0; j < NC; j++)
            {
                if (j >= ncol) break;
                T* ptr = Cs + a.c_n[coff + j];
                const double v = a.alpha * acc[s][j];
                *ptr = a.beta == 0.0 ? T(v) : T(a.beta * double(*ptr) + v);
            }
        }
    }
}

hipError_t launch_combine(int elem_size, const CombineArgs& a, hipStream_t s)
{
    const int64_t nel = a.ms * a.ns;
    if (nel == 0 || a.nsub == 0 || a.nbatch == 0 || a.nprod == 0) return hipSuccess;
    // M in FP32 beside an FP64 C (the FP8 mode's products), else M has C's type
    if (a.m_elem != 0 && a.m_elem != elem_size && !(a.m_elem == 4 && elem_size == 8)) return hipErrorInvalidValue;
    const bool m_f32 = a.m_elem == 4 && elem_size == 8;
    if (a.m_tiled && (!m_f32 || a.m_mp % kMTileRows || a.m_np % kMTileCols || a.m_mp < a.ms || a.m_np < a.ns))
        return hipErrorInvalidValue;
    const int64_t row_tiles = (a.ms + kCombineRows - 1) / kCombineRows;
    const int64_t zb = a.nbatch < 65535 ? a.nbatch : 65535;
    auto blocks_for = [&](int cols_per_lane) {
        return row_tiles * ((a.ns + 4 * cols_per_lane - 1) / (4 * cols_per_lane));
    };
    if (a.dense && (a.nsub == 1 || a.nsub == 4 || a.nsub == 16) && a.nprod <= kCombineMaxProds)
    {
        const int nc = a.nsub == 16 ? 2 : 8;
        const int64_t blocks = blocks_for(nc);
        if (blocks > 0x7fffffff) return hipErrorInvalidValue;
        const dim3 grid(static_cast<unsigned>(blocks), 1u, static_cast<unsigned>(zb));
#define MISTRA_COMBINE_ALL(NSUB, NC)                                                                            \
    do                                                                                                          \
    {                                                                                                           \
        if (m_f32 && a.m_tiled)                                                                                 \
            hipLaunchKernelGGL((combine_all_kernel<double, NSUB, NC, float, true>), grid, dim3(256), 0, s, a);  \
        else if (m_f32)                                                                                         \
            hipLaunchKernelGGL((combine_all_kernel<double, NSUB, NC, float, false>), grid, dim3(256), 0, s, a); \
        else if (elem_size == 8)                                                                                \
            hipLaunchKernelGGL((combine_all_kernel<double, NSUB, NC, double, false>), grid, dim3(256), 0, s, a);\
        else                                                                                                    \
            hipLaunchKernelGGL((combine_all_kernel<float, NSUB, NC, float, false>), grid, dim3(256), 0, s, a);  \
    } while (0)
        if (a.nsub == 16) MISTRA_COMBINE_ALL(16, 2);
        else if (a.nsub == 4) MISTRA_COMBINE_ALL(4, 8);
        else MISTRA_COMBINE_ALL(1, 8);
#undef MISTRA_COMBINE_ALL
        return hipGetLastError();
    }
    const int64_t blocks = blocks_for(kCombineCols);
    if (blocks > 0x7fffffff || a.nsub > 65535) return hipErrorInvalidValue;
    const dim3 grid(static_cast<unsigned>(blocks), static_cast<unsigned>(a.nsub), static_cast<unsigned>(zb));
    if (m_f32 && a.m_tiled) hipLaunchKernelGGL((combine_kernel<double, float, true>), grid, dim3(256), 0, s, a);
    else if (m_f32) hipLaunchKernelGGL((combine_kernel<double, float, false>), grid, dim3(256), 0, s, a);
    else if (elem_size == 8) hipLaunchKernelGGL((combine_kernel<double, double, false>), grid, dim3(256), 0, s, a);
    else hipLaunchKernelGGL((combine_kernel<float, float, false>), grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

// ---- synthetic data ----------------------------------------------------------------------------------

__device__ __forceinline__ uint64_t splitmix64(uint64_t x)
{
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

template <typename T>
__global__ __launch_bounds__(256) void fill_uniform_kernel(T* p, int64_t n, uint64_t seed, int emax)
{
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
    {
        const uint64_t key = (seed << 32) + uint64_t(i);
        double v = double(splitmix64(key) >> 11) * 0x1.0p-53 * 2.0 - 1.0;
        if (emax > 0)
        {
            const uint64_t u = splitmix64(key ^ 0xA5A5A5A5A5A5A5A5ull);
            const int e = int(u % uint64_t(2 * emax + 1)) - emax;
            v = ldexp(v, e);
        }
        p[i] = T(v);
    }
}

hipError_t launch_fill_uniform(int elem_size, void* ptr, int64_t n, uint64_t seed, int emax, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (elem_size == 8)
        hipLaunchKernelGGL(fill_uniform_kernel<double>, dim3(unsigned(blocks)), dim3(256), 0, s,
                           static_cast<double*>(ptr), n, seed, emax);
    else
        hipLaunchKernelGGL(fill_uniform_kernel<float>, dim3(unsigned(blocks)), dim3(256), 0, s,
                           static_cast<float*>(ptr), n, seed, emax);
    return hipGetLastError();
}

// ---- copy ---------------------------------------------------------------------------------------------

__global__ __launch_bounds__(256) void copy16_kernel(uint4* dst, const uint4* src, int64_t n16)
{
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n16; i += int64_t(gridDim.x) * 256)
        dst[i] = src[i];
}

__global__ void copy1_kernel(unsigned char* dst, const unsigned char* src, int64_t n)
{
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

hipError_t launch_copy_bytes(void* dst, const void* src, int64_t nbytes, hipStream_t s)
{
    const int64_t n16 = nbytes / 16;
    if (n16 > 0)
    {
        int64_t blocks = (n16 + 255) / 256;
        if (blocks > 8192) blocks = 8192;
        hipLaunchKernelGGL(copy16_kernel, dim3(unsigned(blocks)), dim3(256), 0, s, static_cast<uint4*>(dst),
                           static_cast<const uint4*>(src), n16);
    }
    const int64_t tail = nbytes - n16 * 16;
    if (tail > 0)
        hipLaunchKernelGGL(copy1_kernel, dim3(1), dim3(64), 0, s, static_cast<unsigned char*>(dst) + n16 * 16,
                           static_cast<const unsigned char*>(src) + n16 * 16, tail);
    return hipGetLastError();
}

// ---- relative error -------------------------------------------------------------------------------------

constexpr int kErrBlocks = 1024;

template <typename T>
__global__ __launch_bounds__(256) void sq_partial_kernel(const T* X, const T* Y, int64_t n, double* part)
{
    __shared__ double s0[256], s1[256];
    double d2 = 0.0, y2 = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
    {
        const double x = double(X[i]), y = double(Y[i]);
        d2 += (x - y) * (x - y);
        y2 += y * y;
    }
    s0[threadIdx.x] = d2;
    s1[threadIdx.x] = y2;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1)
    {
        if (int(threadIdx.x) < w)
        {
            s0[threadIdx.x] += s0[threadIdx.x + w];
            s1[threadIdx.x] += s1[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
    {
        part[2 * blockIdx.x] = s0[0];
        part[2 * blockIdx.x + 1] = s1[0];
    }
}

__global__ __launch_bounds__(256) void sq_final_kernel(const double* part, double* out)
{
    __shared__ double s0[256], s1[256];
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < kErrBlocks; i += 256)
    {
        a += part[2 * i];
        b += part[2 * i + 1];
    }
    s0[threadIdx.x] = a;
    s1[threadIdx.x] = b;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1)
    {
        if (int(threadIdx.x) < w)
        {
            s0[threadIdx.x] += s0[threadIdx.x + w];
            s1[threadIdx.x] += s1[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
    {
        out[0] = s0[0];
        out[1] = s1[0];
    }
}

// `out` must hold 2 + 2*kErrBlocks doubles: the result pair followed by the per-block partials.
hipError_t launch_rel_error(int elem_size, const void* X, const void* Y, int64_t n, double* out, hipStream_t s)
{
    double* part = out + 2;
    if (elem_size == 8)
        hipLaunchKernelGGL(sq_partial_kernel<double>, dim3(kErrBlocks), dim3(256), 0, s,
                           static_cast<const double*>(X), static_cast<const double*>(Y), n, part);
    else
        hipLaunchKernelGGL(sq_partial_kernel<float>, dim3(kErrBlocks), dim3(256), 0, s,
                           static_cast<const float*>(X), static_cast<const float*>(Y), n, part);
    hipLaunchKernelGGL(sq_final_kernel, dim3(1), dim3(256), 0, s, part, out);
    return hipGetLastError();
}

// ---- C = beta*C over a strided tensor ---------------------------------------------------------------------

template <typename T>
__global__ __launch_bounds__(256) void scale_kernel(T* C, StridedDims d, int64_t n, double beta)
{
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
    {
        int64_t rem = i, off = 0;
        for (int k = 0; k < d.ndim; k++)
        {
            off += (rem % d.len[k]) * d.stride[k];
            rem /= d.len[k];
        }
        C[off] = beta == 0.0 ? T(0) : T(beta * double(C[off]));
    }
}

// One workgroup per (64-row x 4-column block, product): every target the product writes first.
__global__ __launch_bounds__(256) void scale_first_targets_kernel(double* C, const int64_t* c_m, const int64_t* c_n,
                                                                  int64_t ms, int64_t ns, const ProdDesc* prods,
                                                                  double beta)
{
    const int64_t row_tiles = (ms + 63) / 64;
    const int64_t r = (int64_t(blockIdx.x) % row_tiles) * 64 + (threadIdx.x & 63);
    const int64_t col = (int64_t(blockIdx.x) / row_tiles) * 4 + (threadIdx.x >> 6);
    if (r >= ms || col >= ns) return;
    const ProdDesc& P = prods[blockIdx.y];
    const int64_t off = c_m[r] + c_n[col];
    for (int t = 0; t < P.nC; t++)
        if ((P.first_mask >> t) & 1)
        {
            double* x = C + P.c[t].base + off;
            *x = beta == 0.0 ? 0.0 : beta * *x;
        }
}

hipError_t launch_scale_first_targets(void* C, const int64_t* c_m, const int64_t* c_n, int64_t ms, int64_t ns,
                                      const ProdDesc* prods, int nprod, double beta, hipStream_t s)
{
    if (ms <= 0 || ns <= 0 || nprod <= 0) return hipSuccess;
    const int64_t blocks = ((ms + 63) / 64) * ((ns + 3) / 4);
    if (blocks > 0x7fffffff || nprod > 65535) return hipErrorInvalidValue;
    hipLaunchKernelGGL(scale_first_targets_kernel, dim3(unsigned(blocks), unsigned(nprod)), dim3(256), 0, s,
                       static_cast<double*>(C), c_m, c_n, ms, ns, prods, beta);
    return hipGetLastError();
}

hipError_t launch_scale(int elem_size, void* C, const StridedDims& d, double beta, hipStream_t s)
{
    int64_t n = 1;
    for (int k = 0; k < d.ndim; k++) n *= d.len[k];
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (elem_size == 8)
        hipLaunchKernelGGL(scale_kernel<double>, dim3(unsigned(blocks)), dim3(256), 0, s, static_cast<double*>(C),
                           d, n, beta);
    else
        hipLaunchKernelGGL(scale_kernel<float>, dim3(unsigned(blocks)), dim3(256), 0, s, static_cast<float*>(C), d,
                           n, beta);
    return hipGetLastError();
}

// ---- FP32 tensors in the low-precision modes: FP64 staging copies -------------------------------------

__global__ __launch_bounds__(256) void f32_to_f64_kernel(const float* __restrict__ src, double* __restrict__ dst,
                                                         int64_t n)
{
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
        dst[i] = double(src[i]);
}

// dst[off] = float(src[off]) for the offsets a strided tensor addresses (elements between them untouched)
__global__ __launch_bounds__(256) void f64_to_f32_strided_kernel(const double* src, float* dst, StridedDims d,
                                                                 int64_t n)
{
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
    {
        int64_t rem = i, off = 0;
        for (int k = 0; k < d.ndim; k++)
        {
            off += (rem % d.len[k]) * d.stride[k];
            rem /= d.len[k];
        }
        dst[off] = float(src[off]);
    }
}

hipError_t launch_f32_to_f64(const float* src, double* dst, int64_t n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(f32_to_f64_kernel, dim3(unsigned(blocks)), dim3(256), 0, s, src, dst, n);
    return hipGetLastError();
}

hipError_t launch_f64_to_f32_strided(const double* src, float* dst, const StridedDims& d, hipStream_t s)
{
    int64_t n = 1;
    for (int k = 0; k < d.ndim; k++) n *= d.len[k];
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(f64_to_f32_strided_kernel, dim3(unsigned(blocks)), dim3(256), 0, s, src, dst, d, n);
    return hipGetLastError();
}

// ---- general index classes -----------------------------------------------------------------------------

template <typename T>
__global__ __launch_bounds__(256) void reduce_kernel(const T* __restrict__ src, T* __restrict__ dst, ReduceArgs a,
                                                     int64_t n_out, int64_t n_red)
{
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n_out; i += int64_t(gridDim.x) * 256)
    {
        int64_t rem = i, base = 0;
        for (int k = 0; k < a.out.ndim; k++)
        {
            base += (rem % a.out.len[k]) * a.out.stride[k];
            rem /= a.out.len[k];
        }
        double sum = 0.0;
        for (int64_t j = 0; j < n_red; j++)
        {
            int64_t r = j, off = base;
            for (int k = 0; k < a.red.ndim; k++)
            {
                off += (r % a.red.len[k]) * a.red.stride[k];
                r /= a.red.len[k];
            }
            sum += double(src[off]);
        }
        dst[i] = T(sum);
    }
}

hipError_t launch_reduce(int elem_size, const void* src, void* dst, const ReduceArgs& a, hipStream_t s)
{
    int64_t n_out = 1, n_red = 1;
    for (int k = 0; k < a.out.ndim; k++) n_out *= a.out.len[k];
    for (int k = 0; k < a.red.ndim; k++) n_red *= a.red.len[k];
    if (n_out <= 0) return hipSuccess;
    int64_t blocks = (n_out + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (elem_size == 8)
        hipLaunchKernelGGL(reduce_kernel<double>, dim3(unsigned(blocks)), dim3(256), 0, s,
                           static_cast<const double*>(src), static_cast<double*>(dst), a, n_out, n_red);
    else
        hipLaunchKernelGGL(reduce_kernel<float>, dim3(unsigned(blocks)), dim3(256), 0, s,
                           static_cast<const float*>(src), static_cast<float*>(dst), a, n_out, n_red);
    return hipGetLastError();
}

template <typename T>
__global__ __launch_bounds__(256) void bcast_add_kernel(const T* __restrict__ R, T* __restrict__ C, BcastArgs a,
                                                        int64_t n)
{
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
    {
        int64_t rem = i, off = 0, roff = 0;
        for (int k = 0; k < a.c.ndim; k++)
        {
            const int64_t x = rem % a.c.len[k];
            rem /= a.c.len[k];
            off += x * a.c.stride[k];
            roff += x * a.r_stride[k];
        }
        const double r = double(R[roff]);
        C[off] = a.beta == 0.0 ? T(r) : T(r + a.beta * double(C[off]));
    }
}

hipError_t launch_bcast_add(int elem_size, const void* R, void* C, const BcastArgs& a, hipStream_t s)
{
    int64_t n = 1;
    for (int k = 0; k < a.c.ndim; k++) n *= a.c.len[k];
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (elem_size == 8)
        hipLaunchKernelGGL(bcast_add_kernel<double>, dim3(unsigned(blocks)), dim3(256), 0, s,
                           static_cast<const double*>(R), static_cast<double*>(C), a, n);
    else
        hipLaunchKernelGGL(bcast_add_kernel<float>, dim3(unsigned(blocks)), dim3(256), 0, s,
                           static_cast<const float*>(R), static_cast<float*>(C), a, n);
    return hipGetLastError();
}

}  // namespace mistra
