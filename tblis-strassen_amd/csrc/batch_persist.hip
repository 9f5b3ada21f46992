// One persistent launch for a batch of contractions (mistra_batch_mult).
//
// The launch works through one device job queue of two kinds of jobs: pack jobs, which write every leaf
// product's operand sums (sum_t c_t A_t, sum_t c_t B_t -- one term for a classical product or a fringe
// pass) as dense K-fast panels, and the GEMM tile jobs of every product.  The host orders the products
// longest remaining chain first and queues each panel's pack jobs well ahead of the product's tiles, so
// memory-bound packing runs beside MFMA-bound tiles on the same CUs.  Every tile job runs the same kernel
// variant: a single-term leaf over K-fast operands with its K-tiles loaded straight into LDS, and the
// multi-C epilogue on the contraction's own C.  A tile waits (bounded) until its product's panels are packed and its contraction's earlier rounds
// are written: the products of one round write disjoint parts of C, rounds that write the same part
// follow each other, as the Strassen stages of one contraction do (src/stra_internal/3t/stra_mult.cxx:
// 445-592 runs the 7 products one after another).  A job waits only for jobs with smaller indices, which
// running workgroups took before it, so the queue always drains.  The last workgroup to leave resets
// the counters for the next launch (a HIP-graph replay needs no memset).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "batch_persist.hpp"
#include "leaf_gemm_kernel.hpp"

#include <map>
#include <mutex>

// Waits of the stage launches that gave up (every launch on the device adds to it; read and cleared by
// the host).  Host code reaches it only through the address hipGetSymbolAddress returns per device.
static __device__ unsigned stage_timeouts;

namespace mistra {

namespace {

// Control flow of the persistent loop is wave-uniform throughout: its counters are read and advanced by
// every lane of wave 0 (adding lane == 0 ? 1 : 0, values broadcast with readfirstlane), never inside a
// "thread 0 only" branch.  A divergent thread-0 block next to the loop's barriers lets the compiler
// merge the block that ends one job with the one that starts the next; wave 0 then reaches the loop's
// first barrier once with lanes 1-63 and once more with lane 0, while the other waves reach it once,
// and the workgroup's barriers fall out of step.

// Wave 0 waits until *c >= target: polls with a short sleep between them, at most spin_limit of them; a
// wait that gives up is counted in timeouts.
__device__ __forceinline__ void wait_count(const unsigned* c, unsigned target, long long spin_limit, unsigned* timeouts,
                                           int lane)
{
    for (long long spins = 0;; spins++)
    {
        const unsigned v = __builtin_amdgcn_readfirstlane(__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        if (v >= target) return;
        if (spins >= spin_limit)
        {
            __hip_atomic_fetch_add(timeouts, lane == 0 ? 1u : 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return;
        }
        __builtin_amdgcn_s_sleep(2);
    }
}

// Wave 0 adds one to *c (one lane's worth) and returns the value before.
__device__ __forceinline__ unsigned bump(unsigned* c, int lane)
{
    return __builtin_amdgcn_readfirstlane(__hip_atomic_fetch_add(c, lane == 0 ? 1u : 0u, __ATOMIC_RELAXED,
                                                                  __HIP_MEMORY_SCOPE_AGENT));
}

// One pack job: rows r0 .. r0 + 31 and column blocks cb0 .. of a packed panel, 32 x 32 at a time:
// out[r][c] = sum_t coef_t X[rs[roff_t + r] + cs[coff_t + c]] for c < cols, 0 for cols <= c < ld.  Sources
// whose unit stride runs along rows are read rows-fastest and transposed through LDS, others are read
// along K; writes run along K, as device-coherent stores (written through to memory: tiles on other
// XCDs read the panel in the same launch).
//
// The job is bound by memory latency, not bandwidth (every element is a scatter-index load followed by a
// data load, and a workgroup holds only four waves), so the column blocks go U at a time: a thread issues
// the index loads of U blocks, then all their data loads, before it uses any of them.  Row offsets are
// loaded once per job.  Positions past the panel read a clamped (valid) element and are zeroed by a
// select, so no load sits behind a divergent branch.
template <typename T>
__device__ __forceinline__ void pack_job(const PackDesc& P, int64_t job, T (*tile)[33])
{
    constexpr int U = sizeof(T) == 8 ? 4 : 8;
    const int64_t cr = (P.rows + 31) / 32, ncb = (P.ld + 31) / 32;
    const int64_t r0 = (job % cr) * 32, cb0 = (job / cr) * kPackColBlocks;
    const int64_t cb1 = cb0 + kPackColBlocks < ncb ? cb0 + kPackColBlocks : ncb;
    const auto X = gptr(static_cast<const T*>(P.X));
    const auto out = gptr(static_cast<T*>(P.out));
    const auto rs = gptr(P.rs);
    const auto cs = gptr(P.cs);
    const int tid = threadIdx.x;
    const int nt = P.nterms;
    auto rclamp = [&](int64_t r) { return r < P.rows ? r : P.rows - 1; };
    auto cclamp = [&](int64_t c) { return c < P.cols ? c : P.cols - 1; };
    T coef[kMaxTerms];
#pragma unroll
    for (int t = 0; t < kMaxTerms; t++) coef[t] = t < nt ? T(P.coef[t]) : T(0);
    if (P.rows_fast)
    {
        // thread: row r0 + (tid % 32), columns c0 + tid / 32 + 8 e of each block
        const int r = tid & 31, cs0 = tid >> 5;
        int64_t roff[kMaxTerms];
#pragma unroll
        for (int t = 0; t < kMaxTerms; t++) roff[t] = t < nt ? rs[P.roff[t] + rclamp(r0 + r)] : 0;
        const bool rin = r0 + r < P.rows;
        for (int64_t cb = cb0; cb < cb1; cb += U)
        {
            T v[U][4];
#pragma unroll
            for (int u = 0; u < U; u++)
#pragma unroll
                for (int e = 0; e < 4; e++) v[u][e] = T(0);
#pragma unroll
            for (int t = 0; t < kMaxTerms; t++)
                if (t < nt)
                {
                    int64_t co[U][4];
#pragma unroll
                    for (int u = 0; u < U; u++)
#pragma unroll
                        for (int e = 0; e < 4; e++) co[u][e] = cs[P.coff[t] + cclamp((cb + u) * 32 + cs0 + 8 * e)];
#pragma unroll
                    for (int u = 0; u < U; u++)
#pragma unroll
                        for (int e = 0; e < 4; e++) v[u][e] += coef[t] * X[roff[t] + co[u][e]];
                }
#pragma unroll
            for (int u = 0; u < U; u++)
                if (cb + u < cb1)
                {
                    const int64_t c0 = (cb + u) * 32;
#pragma unroll
                    for (int e = 0; e < 4; e++)
                    {
                        const int c = cs0 + 8 * e;
                        tile[r][c] = rin && c0 + c < P.cols ? v[u][e] : T(0);
                    }
                    __syncthreads();
#pragma unroll
                    for (int e = 0; e < 4; e++)
                    {
                        const int x = tid + 256 * e, c = x & 31, rr = x >> 5;
                        if (r0 + rr < P.rows && c0 + c < P.ld) coherent_store(&out[(r0 + rr) * P.ld + c0 + c], tile[rr][c]);
                    }
                    __syncthreads();
                }
        }
    }
    else
    {
        // thread: column c0 + (tid % 32) of each block, rows r0 + tid / 32 + 8 e
        const int c = tid & 31, rs0 = tid >> 5;
        int64_t roff[4][kMaxTerms];
#pragma unroll
        for (int t = 0; t < kMaxTerms; t++)
#pragma unroll
            for (int e = 0; e < 4; e++) roff[e][t] = t < nt ? rs[P.roff[t] + rclamp(r0 + rs0 + 8 * e)] : 0;
        for (int64_t cb = cb0; cb < cb1; cb += U)
        {
            T v[U][4];
#pragma unroll
            for (int u = 0; u < U; u++)
#pragma unroll
                for (int e = 0; e < 4; e++) v[u][e] = T(0);
#pragma unroll
            for (int t = 0; t < kMaxTerms; t++)
                if (t < nt)
                {
                    int64_t co[U];
#pragma unroll
                    for (int u = 0; u < U; u++) co[u] = cs[P.coff[t] + cclamp((cb + u) * 32 + c)];
#pragma unroll
                    for (int u = 0; u < U; u++)
#pragma unroll
                        for (int e = 0; e < 4; e++) v[u][e] += coef[t] * X[roff[e][t] + co[u]];
                }
#pragma unroll
            for (int u = 0; u < U; u++)
                if (cb + u < cb1)
                {
                    const int64_t c0 = (cb + u) * 32;
                    const bool cin = c0 + c < P.cols;
#pragma unroll
                    for (int e = 0; e < 4; e++)
                    {
                        const int64_t r = r0 + rs0 + 8 * e;
                        if (r < P.rows && c0 + c < P.ld) coherent_store(&out[r * P.ld + c0 + c], cin ? v[u][e] : T(0));
                    }
                }
        }
    }
}

template <typename V>
__device__ __forceinline__ V uni(V v)
{
    static_assert(sizeof(V) == 8 || sizeof(V) == 4, "scalar fields");
    if constexpr (sizeof(V) == 8)
    {
        int64_t u;
        __builtin_memcpy(&u, &v, 8);
        u = uniform64(u);
        __builtin_memcpy(&v, &u, 8);
    }
    else
    {
        int u;
        __builtin_memcpy(&u, &v, 4);
        u = __builtin_amdgcn_readfirstlane(u);
        __builtin_memcpy(&v, &u, 4);
    }
    return v;
}

// The arguments of one tile job, every field held as wave-uniform: the table is read after a fence, so
// the compiler loads it with vector loads, and buffer resources built from such values would need a
// waterfall loop around every load.
__device__ __forceinline__ GemmArgs uniform_args(const GemmArgs& s)
{
    GemmArgs g;
    g.A = uni(s.A);
    g.B = uni(s.B);
    g.C = uni(s.C);
    g.a_m = uni(s.a_m);
    g.a_k = uni(s.a_k);
    g.a_krun = uni(s.a_krun);
    g.b_k = uni(s.b_k);
    g.b_krun = uni(s.b_krun);
    g.b_n = uni(s.b_n);
    g.c_m = uni(s.c_m);
    g.c_n = uni(s.c_n);
    g.ms = uni(s.ms);
    g.ns = uni(s.ns);
    g.ks = uni(s.ks);
    g.prods = uni(s.prods);
    g.nprod = uni(s.nprod);
    g.tiles_m = uni(s.tiles_m);
    g.tiles_n = uni(s.tiles_n);
    g.alpha = uni(s.alpha);
    g.beta = uni(s.beta);
    g.nbatch = uni(s.nbatch);
    g.batch_a = uni(s.batch_a);
    g.batch_b = uni(s.batch_b);
    g.batch_c = uni(s.batch_c);
    g.M = uni(s.M);
    g.zero = uni(s.zero);
    g.ka_stride = uni(s.ka_stride);
    g.kb_stride = uni(s.kb_stride);
    g.c_scatter = uni(s.c_scatter);
    g.c_coherent = uni(s.c_coherent);
    return g;
}

// Last record e with recs[e].first_job <= j (the records are ordered by first_job; j is the same in every
// lane, so the search runs on the scalar unit).
template <typename R>
__device__ __forceinline__ int find_record(const R* __restrict__ recs, int n, int64_t j)
{
    int lo = 0, hi = n - 1;
    while (lo < hi)
    {
        const int mid = (lo + hi + 1) >> 1;
        if (recs[mid].first_job <= j)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// The tables come as read-only, non-aliased kernel pointers and are indexed with wave-uniform values, so
// their fields load with scalar loads into SGPRs and the leaf body sees its arguments as it does as a
// kernel of its own.
template <typename T, int TM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void batch_persistent_kernel(
    PersistArgs a, const PackDesc* __restrict__ packs, const TileItem* __restrict__ items, const Segment* __restrict__ segs)
{
    constexpr int TK = sizeof(T) == 8 ? 16 : 32;
    __shared__ int64_t s_job;
    __shared__ int s_last;
    __shared__ T tile[32][33];
    unsigned* const queue = a.counters;
    unsigned* const exits = a.counters + 1;
    unsigned* const timeouts = a.counters + 2;
    unsigned* const packed = a.counters + kPersistHeader;          // per product: pack chunks done
    unsigned* const tiles_done = packed + a.nitems;                // per contraction: tile jobs done
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
    for (;;)
    {
        // (the queue is 32-bit: the host keeps jobs plus workgroups below 2^32)
        if (wave == 0) s_job = int64_t(bump(queue, lane));
        __syncthreads();
        const int64_t j = int64_t(__builtin_amdgcn_readfirstlane(unsigned(s_job)));
        __syncthreads();  // every thread has read s_job before wave 0 takes the next job
        if (j >= a.njobs) break;
        const Segment& S = segs[find_record(segs, a.nsegs, j)];
        if (S.kind == 0)
        {
            const PackDesc& P = packs[S.index];
            pack_job<T>(P, j - S.first_job, tile);
            // every thread's panel stores have completed (they are device-coherent) before the count says so
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            if (wave == 0) (void)bump(&packed[P.item], lane);
        }
        else
        {
            const int i = S.index;
            const TileItem& it = items[i];
            if (wave == 0)
            {
                wait_count(&packed[i], unsigned(it.need_pack) + a.dep_bias, a.spin_limit, timeouts, lane);
                if (it.need_entry > 0 || a.dep_bias > 0)
                    wait_count(&tiles_done[it.entry], unsigned(it.need_entry) + a.dep_bias, a.spin_limit, timeouts, lane);
            }
            __syncthreads();
            // Data other XCDs wrote in this launch is read without a cache invalidation: C of a contraction
            // with several rounds goes through device-coherent loads and stores (c_coherent), and a packed
            // panel is written once per launch, after the kernel start has invalidated the caches, in whole
            // 128-byte lines that no other job shares, and read only after its count is complete, so no L2
            // holds an older copy of it.
            const GemmArgs g = uniform_args(it.g);
            leaf_gemm_body<T, TM, TM, TK, 1, false, false, EPI_MULTI_C, KMODE_AFFINE_VEC>(g, int(j - S.first_job), 0, 0);
            // C writes a later round reads (device-coherent stores) have completed before the count says so
            if (it.release) __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            if (wave == 0) (void)bump(&tiles_done[it.entry], lane);
        }
    }
    // the last workgroup out resets the queue and the dependency counters (timeouts stay for the host)
    if (wave == 0) s_last = bump(exits, lane) == gridDim.x - 1;
    __syncthreads();
    if (s_last)
    {
        const int n = a.nitems + a.nentries;
        for (int x = threadIdx.x; x < n; x += 256) packed[x] = 0u;
        if (wave == 0)
        {
            *queue = 0u;
            *exits = 0u;
        }
        __threadfence();
    }
}

// Every stage product of one staged ABC plan in one launch (StageArgs).  g0 holds the arguments the
// products share; a job swaps in its product's table entry.
template <typename T, int TM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void stage_persistent_kernel(
    StageArgs a, GemmArgs g0, const StageProd* __restrict__ prods)
{
    constexpr int TK = sizeof(T) == 8 ? 16 : 32;
    __shared__ int s_pick[3];
    unsigned* const queue = a.counters;
    unsigned* const timeouts = a.timeouts;
    unsigned* const picks = a.counters + kStageHeader;
    unsigned* const tiles = picks + int64_t(kStageXcds) * a.nprod;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
    const int xcd = int(blockIdx.x % kStageXcds);
    const unsigned ntiles = unsigned(a.tiles_m) * unsigned(a.tiles_n);
    const unsigned njobs = ntiles * unsigned(a.nprod);
    for (;;)
    {
        if (wave == 0)
        {
            const unsigned j = bump(queue, lane);
            int prod = -1, tm = 0, tn = 0;
            if (j < njobs)
            {
                prod = int(j / ntiles);
                // a tile of this product: from this XCD's block, else from the next block with tiles left
                // (the product has as many jobs as tiles, so some block always has one)
                for (int k = 0; k < kStageXcds; k++)
                {
                    const int x = (xcd + k) % kStageXcds;
                    const int r0 = a.rb[x / a.pc], r1 = a.rb[x / a.pc + 1];
                    const int c0 = a.cb[x % a.pc], c1 = a.cb[x % a.pc + 1];
                    const unsigned len = unsigned(r1 - r0) * unsigned(c1 - c0);
                    if (len == 0) continue;
                    const unsigned v = bump(&picks[int64_t(prod) * kStageXcds + x], lane);
                    if (v < len)
                    {
                        tm = r0 + int(v / unsigned(c1 - c0));
                        tn = c0 + int(v % unsigned(c1 - c0));
                        break;
                    }
                }
                const StageProd& P = prods[prod];
                const int64_t at = int64_t(tm) * a.tiles_n + tn;
                for (int c = 0; c < P.nC; c++)
                    wait_count(&tiles[P.cbase[c] + at], P.need[c] + a.dep_bias, a.spin_limit, timeouts, lane);
            }
            s_pick[0] = prod;
            s_pick[1] = tm;
            s_pick[2] = tn;
        }
        __syncthreads();
        const int prod = __builtin_amdgcn_readfirstlane(s_pick[0]);
        const int tm = __builtin_amdgcn_readfirstlane(s_pick[1]);
        const int tn = __builtin_amdgcn_readfirstlane(s_pick[2]);
        __syncthreads();  // every thread has read the pick before wave 0 takes the next job
        if (prod < 0) break;
        const StageProd& P = prods[prod];
        GemmArgs g = g0;
        g.prods = g0.prods + prod;
        leaf_gemm_body<T, TM, TM, TK, 1, false, false, EPI_MULTI_C, KMODE_AFFINE_VEC>(g, 0, 0, 0, tm, tn);
        // C stores a later product reads (device-coherent stores) have completed before the counts say so
        if (P.release) __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (wave == 0)
        {
            const int64_t at = int64_t(tm) * a.tiles_n + tn;
            for (int c = 0; c < P.nC; c++) (void)bump(&tiles[P.cbase[c] + at], lane);
        }
    }
}

__global__ void zero_u32_kernel(unsigned* p, int64_t n)
{
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) p[i] = 0u;
}

template <typename T, int TM>
hipError_t launch_stage_t(const StageArgs& a, const GemmArgs& g, int grid, hipStream_t s)
{
    hipLaunchKernelGGL((stage_persistent_kernel<T, TM>), dim3(unsigned(grid)), dim3(256), 0, s, a, g, a.prods);
    return hipGetLastError();
}

template <typename T, int TM>
int stage_occupancy_t()
{
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, reinterpret_cast<const void*>(&stage_persistent_kernel<T, TM>),
                                                     256, 0) != hipSuccess)
    {
        (void)hipGetLastError();
        return 1;
    }
    return n > 0 ? n : 1;
}

template <typename T, int TM>
hipError_t launch_t(const PersistArgs& a, int grid, hipStream_t s)
{
    hipLaunchKernelGGL((batch_persistent_kernel<T, TM>), dim3(unsigned(grid)), dim3(256), 0, s, a, a.packs, a.items, a.segs);
    return hipGetLastError();
}

template <typename T, int TM>
int occupancy_t()
{
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, reinterpret_cast<const void*>(&batch_persistent_kernel<T, TM>),
                                                     256, 0) != hipSuccess)
    {
        (void)hipGetLastError();
        return 1;
    }
    return n > 0 ? n : 1;
}

}  // namespace

hipError_t launch_batch_persistent(int elem, int tile, const PersistArgs& a, int grid, hipStream_t s)
{
    if (elem == 8) return tile == TILE_64 ? launch_t<double, 64>(a, grid, s) : launch_t<double, 128>(a, grid, s);
    return tile == TILE_64 ? launch_t<float, 64>(a, grid, s) : launch_t<float, 128>(a, grid, s);
}

hipError_t launch_stage_persistent(int elem, int tile, const StageArgs& a, const GemmArgs& g, int grid, hipStream_t s)
{
    if (elem == 8) return tile == TILE_64 ? launch_stage_t<double, 64>(a, g, grid, s) : launch_stage_t<double, 128>(a, g, grid, s);
    return tile == TILE_64 ? launch_stage_t<float, 64>(a, g, grid, s) : launch_stage_t<float, 128>(a, g, grid, s);
}

int stage_persistent_occupancy(int elem, int tile)
{
    if (elem == 8) return tile == TILE_64 ? stage_occupancy_t<double, 64>() : stage_occupancy_t<double, 128>();
    return tile == TILE_64 ? stage_occupancy_t<float, 64>() : stage_occupancy_t<float, 128>();
}

hipError_t stage_timeouts_word(unsigned** p)
{
    static std::mutex mu;
    static std::map<int, unsigned*> at;  // device -> address of its copy of stage_timeouts
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lk(mu);
    auto it = at.find(dev);
    if (it == at.end())
    {
        void* q = nullptr;
        e = hipGetSymbolAddress(&q, HIP_SYMBOL(stage_timeouts));
        if (e != hipSuccess) return e;
        it = at.emplace(dev, static_cast<unsigned*>(q)).first;
    }
    *p = it->second;
    return hipSuccess;
}

hipError_t stage_timeouts_take(unsigned* n)
{
    unsigned* p = nullptr;
    hipError_t e = stage_timeouts_word(&p);
    if (e == hipSuccess) e = hipMemcpy(n, p, sizeof(unsigned), hipMemcpyDeviceToHost);
    if (e == hipSuccess && *n)
    {
        const unsigned zero = 0;
        e = hipMemcpy(p, &zero, sizeof(unsigned), hipMemcpyHostToDevice);
    }
    return e;
}

hipError_t launch_zero_u32(unsigned* p, int64_t n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 1024);
    hipLaunchKernelGGL(zero_u32_kernel, dim3(unsigned(blocks)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

int batch_persistent_occupancy(int elem, int tile)
{
    if (elem == 8) return tile == TILE_64 ? occupancy_t<double, 64>() : occupancy_t<double, 128>();
    return tile == TILE_64 ? occupancy_t<float, 64>() : occupancy_t<float, 128>();
}

}  // namespace mistra
