// Descriptors of the persistent batch launch (batch_persist.hip), shared with the host (batch.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace mistra {

// One operand panel of one leaf product: out[r][c] (rows x ld, K fastest, zero for cols <= c < ld) =
// sum_t coef_t X[rs[roff_t + r] + cs[coff_t + c]].  rows = M (A) or N (B), cols = K.
struct PackDesc
{
    const void* X;
    void* out;
    const int64_t* rs;
    const int64_t* cs;
    int64_t rows, cols, ld;
    int64_t roff[kMaxTerms], coff[kMaxTerms];
    double coef[kMaxTerms];
    int nterms;
    int rows_fast;      // the source's unit stride runs along rows (transposed through LDS)
    int item;           // the product whose pack counter this panel's chunks advance
    int64_t first_job;  // queue index of its first pack job (pack_jobs_of(rows, ld) of them)
};

// One leaf product (or fringe pass) of one contraction over packed panels: tile jobs
// [first_job, first_job + tiles_m * tiles_n) of the tile range.
struct TileItem
{
    GemmArgs g;
    int entry;
    int round;
    int64_t first_job;
    int64_t need_pack;   // pack jobs of its two panels
    int64_t need_entry;  // tile jobs of its contraction's earlier rounds
    int release;         // a later round of its contraction reads its C writes (completed before counting)
};

// A pack job covers 32 rows and up to kPackColBlocks blocks of 32 columns of one panel (the fence and the
// counter update at its end are paid once per job, not per 32 x 32 block).
constexpr int kPackColBlocks = 32;
__host__ __device__ inline int64_t pack_jobs_of(int64_t rows, int64_t ld)
{
    return ((rows + 31) / 32) * ((((ld + 31) / 32) + kPackColBlocks - 1) / kPackColBlocks);
}

// counters: [0] queue, [1] exit tickets, [2] timed-out waits, [3] unused; then nitems pack counters and
// nentries tile counters (all zero between launches; the last workgroup of a launch resets them)
constexpr int kPersistHeader = 4;

struct PersistArgs
{
    const PackDesc* packs;
    int npacks;
    const TileItem* items;
    int nitems;
    int nentries;
    int64_t njobs_pack, njobs;  // pack jobs [0, njobs_pack), tile jobs [njobs_pack, njobs)
    unsigned* counters;
    long long spin_limit;       // polls a wait makes before it gives up (and counts a timeout)
    unsigned dep_bias;          // added to every wait's target (test hook: 0 in normal use)
};

hipError_t launch_batch_persistent(int elem, int tile, const PersistArgs& a, int grid, hipStream_t s);
// resident workgroups per CU of the persistent kernel
int batch_persistent_occupancy(int elem, int tile);

}  // namespace mistra
