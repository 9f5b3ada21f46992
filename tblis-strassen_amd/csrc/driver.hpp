// Host driver of the Strassen tensor contraction: plans (device-resident scatter vectors and product
// tables, built once per contraction shape) and their execution on a HIP stream.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "batch_persist.hpp"
#include "kernels.hpp"
#include "launch.hpp"
#include "plan.hpp"

namespace mistra {

// Status carrier: code 0 = ok, else a MISTRA_ERR_* code with a message.
struct Status
{
    int code = 0;
    std::string msg;
    bool ok() const { return code == 0; }
    static Status error(int c, std::string m) { return Status{c, std::move(m)}; }
};

// Implementation choice, numbered as tblis::internal::impl_t.
enum Impl
{
    IMPL_ABC = 0,
    IMPL_BLAS = 1,
    IMPL_REF = 2,
    IMPL_NAIVE = 3,
    IMPL_AB = 4
};

// Fast-K eligibility of one launch (see kfast_check in driver.cpp).
struct KFast
{
    bool on = false;          // a fast-body mode (the launch uses the shifted product table)
    int mode = KMODE_GENERAL;  // KMode of the launch
    int64_t sa = 0, sb = 0;   // common K-run strides of A and B (KMODE_AFFINE)
};

struct Pass  // one classical launch over a sub-range (the fringe the 2^L split leaves over)
{
    int64_t ms, ns, ks;
    int prod_index;  // into the device product array
    KFast kf;
};

struct Plan
{
    int elem = 8;  // element size in bytes
    int level = 0;
    int impl = IMPL_ABC;
    int tile = TILE_128;
    bool tile_given = false;  // the tile came from the caller or the autotuner (also for Naive's dense GEMM)
    unsigned flags = 0;
    int64_t m = 0, n = 0, k = 0, nbatch = 1;
    int64_t d = 1, ms = 0, ns = 0, ks = 0;  // split factor and leaf extents
    bool transposed = false;
    bool a_mfast = true, b_nfast = true;
    bool general = false;   // A-only / B-only / C-only indices present (handled by the reduce path)
    bool core = false;      // Strassen core (level >= 1 with non-empty leaves) present

    // host copies of the index data (kept for the reduce path and for tests)
    std::vector<int64_t> h_a_m, h_a_k, h_b_k, h_b_n, h_c_m, h_c_n, h_batch_a, h_batch_b, h_batch_c;
    // step d when the leaf's K offsets are affine (k_off[k] = k_off[0] + k d for k < ks), else -1
    int64_t a_k_step = -1, b_k_step = -1;
    std::vector<ProdDesc> h_prods;  // [0, nprod) core products, then one per fringe pass
    int nprod = 0;
    std::vector<Pass> fringe;

    // combine tables (AB / Naive)
    int nsub = 0;
    std::vector<int64_t> h_sub_roff, h_sub_coff;
    std::vector<int> h_term_begin;
    std::vector<CombineTerm> h_terms;
    std::vector<unsigned> h_comb_mask;  // the same as CombineArgs::pmask (empty if not representable)

    // Naive: dense products over S_p (ms x ks) and T_p (ks x ns)
    std::vector<ProdDesc> h_dense_prods;

    // ABC launch stages: products [stage_begin[s], stage_begin[s + 1]) of the staged tables run in one
    // launch; their C targets are disjoint.  Every C sub-block sees its products in stage order, and
    // h_stage_first holds the first-writer masks (beta) for that order, per staged product.
    std::vector<int> h_stage_begin, h_stage_terms, h_stage_first;
    std::vector<int> h_stage_order;  // product index of each staged position (make_stages' order)
    const ProdDesc* prods_staged = nullptr;
    const ProdDesc* prods_fast_staged = nullptr;

    // fast-K eligibility of the core launch and of the Naive dense launch
    KFast kf_core, kf_dense;

    // ABC with operand sums staged: the sums of the first staged_depth levels (7 or 49 of them) of A and
    // of B are written as dense K-fast buffers S (n x os_m x os_k) and T (n x os_k x os_n), and the stage
    // launches read the inner sums of those (2 terms when only the outer level of 2 is staged, else a
    // single term) instead of 4- (2-) term sums of A and B.  C is still updated in the leaf kernel's epilogue, in the same stages and order as the
    // fused form.
    bool outer_staged = false;
    int staged_depth = 0;                   // levels whose sums are staged (1: the outer 7; 2: all 49)
    int64_t os_m = 0, os_n = 0, os_k = 0;  // staged sub-block extents (m / 2^depth, ... rounded to leaves)
    std::vector<ProdDesc> h_outer_sums;    // sum-kernel tables: A with rows = K, cols = M; B rows = K, cols = N
    std::vector<ProdDesc> h_outer_shifted; // the same with term bases = shifts from the first sub-block
    bool outer_shift_a = false, outer_shift_b = false;  // every outer term of A (B) is a translate
    bool outer_pair_a = false, outer_pair_b = false;    // rows in contiguous, even-aligned pairs
    // one-pass sums (launch_operand_multisums) when every source sub-block of A (B) is a translate:
    // coefficient matrix [nout][nsrc] over the sub-blocks of the staged split, and their shifts
    bool msum_a = false, msum_b = false;
    // the one-pass sums may read two K rows per load: K comes in contiguous pairs at even offsets, and
    // every column offset and source shift is even (the batch offset and base address are checked at launch)
    bool msum_pair_a = false, msum_pair_b = false;
    int msum_nsrc = 0;
    std::vector<double> h_msum_coef_a, h_msum_coef_b;
    std::vector<double> h_msum_coef1_a, h_msum_coef1_b;  // 1-level 7 x 4 matrices (Kronecker factor, depth 2)
    const double *msum_coef1_a = nullptr, *msum_coef1_b = nullptr;
    std::vector<int64_t> h_msum_shift_a, h_msum_shift_b;
    const double *msum_coef_a = nullptr, *msum_coef_b = nullptr;
    const int64_t *msum_shift_a = nullptr, *msum_shift_b = nullptr;
    std::vector<ProdDesc> h_inner;         // the products over S / T in table order (fast-body form)
    std::vector<int> h_inner_stage_terms;
    KFast kf_inner;
    const ProdDesc* outer_sums = nullptr;
    const ProdDesc* outer_shifted = nullptr;
    const ProdDesc* inner_staged = nullptr;
    const int64_t *os_sm = nullptr, *os_sk = nullptr, *os_skrun = nullptr, *os_tk = nullptr, *os_tkrun = nullptr,
                  *os_tn = nullptr;
    // The staged products in one persistent launch (stage_persistent_kernel, MISTRA_FLAG_STAGE_PERSIST)
    // instead of one launch per stage: when every staged product has one term per operand and whole
    // K-tiles.  h_stage_args holds
    // everything but the per-call counters; stage_targets is the number of distinct C targets.
    bool stage_persist = false;
    std::vector<ProdDesc> h_stage_pdesc;   // the staged products in launch order (first writers for it)
    const ProdDesc* stage_pdesc = nullptr;
    std::vector<StageProd> h_stage_prods;
    StageArgs h_stage_args{};
    int stage_targets = 0;

    // device mirror (one allocation)
    void* dmem = nullptr;
    const int64_t *a_m = nullptr, *a_k = nullptr, *a_krun = nullptr, *b_k = nullptr, *b_krun = nullptr,
                  *b_n = nullptr, *c_m = nullptr, *c_n = nullptr;
    const int64_t *batch_a = nullptr, *batch_b = nullptr, *batch_c = nullptr;
    const ProdDesc* prods = nullptr;
    const ProdDesc* prods_fast = nullptr;  // bases = term shifts (fast kernel body)
    const int64_t *sub_roff = nullptr, *sub_coff = nullptr;
    const int* term_begin = nullptr;
    const CombineTerm* terms = nullptr;
    const int64_t *s_m = nullptr, *s_k = nullptr, *s_krun = nullptr, *t_k = nullptr, *t_krun = nullptr,
                  *t_n = nullptr;
    const ProdDesc* dense_prods = nullptr;
    // Naive with S written K-fast (one-pass sums): row offsets i * ks of S; its K vector is t_k
    const int64_t* sk_m = nullptr;
    KFast kf_dense_k;
    const void* zero = nullptr;  // 256 zeroed bytes

    // general (non-pure) contractions: reduce/broadcast descriptors
    Grouping grouping;

    ~Plan();
};

// Device workspace handed out from one fixed block (bump allocation, nothing freed) instead of the
// stream-ordered pool: callers that must not allocate pass one.
struct Arena
{
    char* base = nullptr;
    size_t size = 0;
    size_t used = 0;
};

struct Call
{
    const void* A;
    const void* B;
    void* C;
    double alpha, beta;
    Arena* ws = nullptr;  // null: workspace from hipMallocAsync; else from this arena (and no side stream)
    unsigned flags = 0;   // the call's MISTRA_FLAG_* (MISTRA_FLAG_STAGE_PERSIST is read at run time)
};

// Device image of a plan: every table in one byte block (256-byte aligned parts) and the offset of each.
struct PlanImage
{
    static constexpr int kParts = 45;
    std::vector<unsigned char> bytes;
    size_t off[kParts];
};
void build_image(Plan& p, PlanImage& img);
void bind_image(Plan& p, const PlanImage& img, void* base);

// Build (or fetch from the cache) the plan of a contraction.  capturing: the call is being recorded
// into a HIP graph; then the plan must already exist (an error otherwise) and is pinned for the life of
// the process, since the graph holds its device table addresses.
Status get_plan(const tblis::tblis_tensor* A, const label_type* idx_A, const tblis::tblis_tensor* B,
                const label_type* idx_B, const tblis::tblis_tensor* C, const label_type* idx_C, int level,
                int impl, int tile, unsigned flags, std::shared_ptr<Plan>& out, bool capturing = false);

// Cache key of a plan: device, type, level, impl, tile, the flags that change tables, and every tensor's
// labels, lengths and strides.
std::string plan_key(const tblis::tblis_tensor* A, const label_type* idx_A, const tblis::tblis_tensor* B,
                     const label_type* idx_B, const tblis::tblis_tensor* C, const label_type* idx_C, int level,
                     int impl, int tile, unsigned flags, int dev);

// After the stream has drained: an error if a persistent stage launch on this device gave up a wait
// since the last check (its results are then not valid); clears the count.
Status stage_wait_status();

// The allocation-free entry point's device scratch block on the current device (noalloc.cpp): its
// address (from hipGetSymbolAddress, looked up once per device) and size.
Status noalloc_scratch(void** ptr, size_t* bytes);

// Build a plan without touching the device (index analysis and tables only; for tests and tuning).
Status build_plan_host(const tblis::tblis_tensor* A, const label_type* idx_A, const tblis::tblis_tensor* B,
                       const label_type* idx_B, const tblis::tblis_tensor* C, const label_type* idx_C,
                       int level, int impl, int tile, unsigned flags, Plan& p);

// ABC launch stages of a plan's core products (fills h_stage_begin / h_stage_terms; `order` = the
// product indices in stage order).
void make_stages(Plan& p, std::vector<int>& order);

// One multi-C launch of an ABC plan, as the batch path groups them: launches of equal `round` may run
// together (they write disjoint C); rounds run in order.
struct LaunchJob
{
    GemmArgs g;  // tiles_m / tiles_n filled for the effective tile
    int elem, tile, terms, round;
    int kmode;
    bool a_mfast, b_nfast;
    double flops;
};
Status plan_abc_jobs(const Plan& p, const Call& c, std::vector<LaunchJob>& out);

// Execute a plan on device pointers.
Status run_plan(const Plan& p, const Call& c, hipStream_t stream);
Status run_fringe(const Plan& p, const Call& c, hipStream_t stream);
// Low-precision leaf products (MISTRA_PREC_*) of an FP64 plan built with IMPL_NAIVE tables.
Status run_plan_lowp(const Plan& p, const Call& c, int precision, int nslices, hipStream_t stream);

// Whole-call entry used by every C-ABI symbol.  level < 0 or impl == MISTRA_IMPL_AUTO: the autotuner's
// choice (a measured one when mistra_autotune has timed this shape, else the analytic model); tile < 0:
// the plan's default register tile.
Status tensor_mult(const tblis::tblis_tensor* A, const label_type* idx_A, const tblis::tblis_tensor* B,
                   const label_type* idx_B, tblis::tblis_tensor* C, const label_type* idx_C, int level, int impl,
                   int precision, unsigned flags, hipStream_t stream, bool synchronous, int tile = -1);

// Autotuner (autotune.cpp): level, impl and register tile for a contraction shape.
struct TuneChoice
{
    int level = 1, impl = IMPL_ABC, tile = -1;
    double ms = 0.0;        // measured (or modelled) milliseconds per call
    bool measured = false;
};
Status autotune(const tblis::tblis_tensor* A, const label_type* idx_A, const tblis::tblis_tensor* B,
                const label_type* idx_B, const tblis::tblis_tensor* C, const label_type* idx_C, bool measure,
                hipStream_t stream, TuneChoice& out);
// The modelled best level for a given impl (host only).
int best_level(const tblis::tblis_tensor* A, const label_type* idx_A, const tblis::tblis_tensor* B,
               const label_type* idx_B, const tblis::tblis_tensor* C, const label_type* idx_C, int impl);
// Modelled milliseconds of one candidate (host only).
double model_ms(const Plan& p);
// MISTRA_TILE (tuning): the register tile every plan is forced to, or -1
int tile_override();
// Whether a plan makes a single-term leaf launch with affine K addressing, the only launches the
// rectangular tiles serve.
bool has_single_term_launch(const Plan& p);
// Every (level, impl, tile) the autotuner considers for a shape, with its modelled time, fastest first;
// `general`: the contraction takes the general path (no candidates).
struct TuneCand
{
    int level, impl, tile;
    double ms;
};
Status tune_candidates(const tblis::tblis_tensor* A, const label_type* idx_A, const tblis::tblis_tensor* B,
                       const label_type* idx_B, const tblis::tblis_tensor* C, const label_type* idx_C,
                       std::vector<TuneCand>& cands, bool& general);

// Batched contractions (batch.cpp).
Status batch_mult(int count, const mistra_contraction* entries, unsigned flags, hipStream_t stream);
// The persistent launch's tables for these entries, built on the host alone and checked by replaying the
// job queue in order (no device work; for tests).
Status batch_check_host(int count, const mistra_contraction* entries);
void release_batch_graphs();

// 1: the combine pass always takes its per-sub-block form (mistra_debug_set_combine_form)
extern std::atomic<int> g_combine_per_sub;
// test overrides of the Ozaki modes (mistra_debug_set_lowp; 0 = default): BF16 and INT8 digit counts,
// and the INT8 GEMM's fold period in K-steps
extern std::atomic<int> g_dbg_ozaki_slices, g_dbg_i8_slices, g_dbg_i8_fold;
// register tile of every later plan built without an explicit one (mistra_debug_set_tile; -1 = off)
extern std::atomic<int> g_dbg_tile;
// 1: low-precision plans stage their operand sums before slicing them, even where one launch could do both
// (mistra_debug_set_lowp_fusion)
extern std::atomic<int> g_dbg_unfused_lowp;
double scalar_value(const tblis::tblis_tensor* t);
void set_scalar(tblis::tblis_tensor* t, double v);
// test overrides of the persistent launches' waits (mistra_debug_set_batch_wait)
extern std::atomic<long long> g_dbg_batch_spin;
extern std::atomic<unsigned> g_dbg_batch_bias;

// Kernel timing hook of the dominant leaf kernel.
void timing_begin(hipStream_t s);
void timing_end(hipStream_t s, double flops);

void release_all();

// Stream-ordered workspace (hipMallocAsync on the device's default pool).  The pool keeps what calls
// free rather than handing it back to the driver at every synchronization (its default release
// threshold is 0): a call after a synchronization would otherwise map its workspace again (about 40 ms
// for the 3.5 GiB of an Ozaki BF16 1-level call at 8192^3).  trim_stream_pool() gives it back
// (mistra_release_workspace).
hipError_t stream_alloc(void** out, size_t bytes, hipStream_t stream);
void trim_stream_pool();

}  // namespace mistra
