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
    int64_t first_job;  // queue position of its first pack job (pack_jobs_of(rows, ld) of them)
};

// One leaf product (or fringe pass) of one contraction over packed panels: tile jobs at queue positions
// [first_job, first_job + tiles_m * tiles_n).
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

// A run of consecutive queue positions: the pack jobs of one panel (kind 0, index into the pack table) or
// the tile jobs of one product (kind 1, index into the item table).  Segments are stored in queue order.
struct Segment
{
    int64_t first_job;
    int kind;
    int index;
};

// counters: [0] queue, [1] exit tickets, [2] timed-out waits, [3] unused; then nitems pack counters and
// nentries tile counters (all zero between launches; the last workgroup of a launch resets them)
constexpr int kPersistHeader = 4;

struct PersistArgs
{
    const PackDesc* packs;
    int npacks;
    const TileItem* items;
    int nitems;
    const Segment* segs;
    int nsegs;
    int nentries;
    int64_t njobs;              // queue positions [0, njobs), covered by the segments
    unsigned* counters;
    long long spin_limit;       // polls a wait makes before it gives up (and counts a timeout)
    unsigned dep_bias;          // added to every wait's target (test hook: 0 in normal use)
};

// ---- one staged ABC contraction: every stage product in one persistent launch -------------------------
//
// The products (in stage order) all run over the same tile grid with the same arguments but the product
// table entry.  Job j of the queue is one tile of product j / tiles; the workgroup that takes it picks
// which tile through per-(product, block) counters, its own XCD's block of the grid first (so the XCD's
// L2 serves the operand panels its tiles share), then the other blocks.  A tile waits until every
// earlier product that updates one of its C targets has finished the same tile of that target.

constexpr int kStageXcds = 8;  // XCDs of the device: workgroup b runs on XCD b % 8 (round-robin dispatch)

// One product's dependencies: per C target, the counter block of that target (one counter per tile,
// row-major over the grid) and how many earlier products update the target.
struct StageProd
{
    int64_t cbase[kMaxTerms];
    unsigned need[kMaxTerms];
    int nC;
    int release;  // a later product reads one of its targets (its C stores complete before it counts)
};

// counters: [0] queue, then kStageXcds per product (tiles picked from each block), then the per-target
// tile counters; all zero at launch.
constexpr int kStageHeader = 1;

struct StageArgs
{
    const StageProd* prods;
    int nprod;
    int tiles_m, tiles_n;
    int pr, pc;                  // the grid's blocks: pr x pc = kStageXcds, block x = (x / pc, x % pc)
    int rb[kStageXcds + 1], cb[kStageXcds + 1];  // block row bounds rb[0..pr], column bounds cb[0..pc]
    unsigned* counters;
    unsigned* timeouts;          // waits that gave up are counted here
    long long spin_limit;
    unsigned dep_bias;           // added to every wait's target (test hook: 0 in normal use)
};

// Counters a plan's stage launch needs (header, pick counters, tile counters of ntargets targets).
inline int64_t stage_counters(int nprod, int ntargets, int tiles_m, int tiles_n)
{
    return kStageHeader + int64_t(kStageXcds) * nprod + int64_t(ntargets) * tiles_m * tiles_n;
}

hipError_t launch_stage_persistent(int elem, int tile, const StageArgs& a, const GemmArgs& g, int grid, hipStream_t s);
int stage_persistent_occupancy(int elem, int tile);
hipError_t launch_zero_u32(unsigned* p, int64_t n, hipStream_t s);
// The current device's count of stage-launch waits that gave up (a device word shared by every launch
// on the device): its address, and its value read and cleared (after the stream has drained).
hipError_t stage_timeouts_word(unsigned** p);
hipError_t stage_timeouts_take(unsigned* n);

hipError_t launch_batch_persistent(int elem, int tile, const PersistArgs& a, int grid, hipStream_t s);
// resident workgroups per CU of the persistent kernel
int batch_persistent_occupancy(int elem, int tile);

}  // namespace mistra
