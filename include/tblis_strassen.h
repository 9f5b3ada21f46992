/*
 * tblis_strassen.h -- C-ABI of the MI355X (gfx950) Strassen tensor contraction library.
 *
 * Drop-in surface for the Strassen tensor-contraction path of flame/tblis-strassen:
 *   - the tblis_tensor struct, byte-identical to src/util/basic_types.h:341-419
 *     (len_type = stride_type = ptrdiff_t, label_type = char, the configure defaults of
 *     configure.ac:77-103);
 *   - stra_tblis_tensor_mult        replaces src/stra_iface/3t/stra_mult.h:17-20       (1-level Strassen)
 *   - stra_tblis_tensor_mult_2level replaces src/stra_iface/3t/stra_mult_2level.h:17-20 (2-level Strassen)
 *   - tblis_tensor_mult             replaces src/iface/3t/mult.h:17-20                 (classical, L=0)
 * The three are declared extern "C" inside namespace tblis, as in the reference, so their unmangled
 * names are the ABI.  Semantics: C = alpha_A*alpha_B * A.B + alpha_C * C, with alpha in tblis_tensor.scalar;
 * on return C->scalar = 1 and C->conj = 0 (src/stra_iface/3t/stra_mult.cxx:251-252).
 *
 * Tensor data may live in device memory (HBM, the fast path) or in host memory (staged through HBM).
 * The impl (ABC / AB / Naive) of the drop-in symbols is the process-global tblis::internal::impl
 * (src/internal/3t/mult.cxx:26); C callers use mistra_set_impl().
 *
 * Everything prefixed mistra_ is an extension without a reference counterpart: explicit level, impl,
 * precision and HIP stream, error codes instead of void, batching, autotuning and adjoints.
 */
#ifndef TBLIS_STRASSEN_AMD_H
#define TBLIS_STRASSEN_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
namespace tblis {
extern "C" {
#endif

typedef ptrdiff_t len_type;     /* src/util/basic_types.h:61, configure default */
typedef ptrdiff_t stride_type;  /* src/util/basic_types.h:62, configure default */
typedef char label_type;        /* src/util/basic_types.h:63, configure default */

/* src/util/basic_types.h:52-58 */
typedef enum
{
    TYPE_SINGLE   = 0,
    TYPE_FLOAT    = TYPE_SINGLE,
    TYPE_DOUBLE   = 1,
    TYPE_SCOMPLEX = 2,
    TYPE_DCOMPLEX = 3
} type_t;

/* Opaque in this library: a thread communicator (tci_comm) and a CPU kernel configuration in the
 * reference.  Both are accepted for ABI compatibility and ignored: the GPU grid replaces the thread
 * team and the kernel table replaces the CPU configuration. */
typedef struct tci_comm tblis_comm;
typedef struct tblis_config_s tblis_config;

/* src/util/basic_types.h:341-419 (same member order, alignment and size: 56 bytes on LP64) */
typedef struct tblis_tensor
{
    type_t type;
    int conj;
    char scalar[16] __attribute__((__aligned__(8)));
    void* data;
    unsigned ndim;
    len_type* len;
    stride_type* stride;
} tblis_tensor;

/* src/util/basic_types.h:264-272 (same member order and alignment: 64 bytes on LP64). */
typedef struct tblis_matrix
{
    type_t type;
    int conj;
    char scalar[16] __attribute__((__aligned__(8)));
    void* data;
    len_type m, n;
    stride_type rs, cs;
} tblis_matrix;

/* src/util/thread.h:8 -- a one-thread communicator; here only a distinguishable non-null pointer. */
extern const tblis_comm* const tblis_single;

/* src/stra_iface/3t/stra_mult.h:17-20: one-level Strassen tensor contraction. */
void stra_tblis_tensor_mult(const tblis_comm* comm, const tblis_config* cfg,
                            const tblis_tensor* A, const label_type* idx_A,
                            const tblis_tensor* B, const label_type* idx_B,
                                  tblis_tensor* C, const label_type* idx_C);

/* src/stra_iface/3t/stra_mult_2level.h:17-20: two-level Strassen tensor contraction. */
void stra_tblis_tensor_mult_2level(const tblis_comm* comm, const tblis_config* cfg,
                                   const tblis_tensor* A, const label_type* idx_A,
                                   const tblis_tensor* B, const label_type* idx_B,
                                         tblis_tensor* C, const label_type* idx_C);

/* src/iface/3t/mult.h:17-20: classical (L=0) tensor contraction. */
void tblis_tensor_mult(const tblis_comm* comm, const tblis_config* cfg,
                       const tblis_tensor* A, const label_type* idx_A,
                       const tblis_tensor* B, const label_type* idx_B,
                             tblis_tensor* C, const label_type* idx_C);

/* ------------------------------------------------------------------------------------------------
 * Extensions (no reference counterpart).
 * ------------------------------------------------------------------------------------------------ */

/* Implementation selector; numeric values follow tblis::internal::impl_t (src/internal/3t/mult.hpp:14). */
enum
{
    MISTRA_IMPL_ABC   = 0,  /* BLIS_BASED: sums fused into the LDS gather, multi-C epilogue   */
    MISTRA_IMPL_BLAS  = 1,  /* BLAS_BASED: treated as ABC on the GPU                         */
    MISTRA_IMPL_REF   = 2,  /* REFERENCE: brute-force classical contraction kernel           */
    MISTRA_IMPL_NAIVE = 3,  /* STRA_NAIVE: explicit S=sum(A), T=sum(B), plain GEMM, combine   */
    MISTRA_IMPL_AB    = 4,  /* STRA_AB: fused sums, products staged in HBM, one combine pass  */
    MISTRA_IMPL_AUTO  = -1  /* let the autotuner choose level, impl and tile shape            */
};

/* Arithmetic of the leaf products. */
enum
{
    MISTRA_PREC_NATIVE     = 0, /* FP64 MFMA for double, FP32 MFMA for float                          */
    MISTRA_PREC_BF16X3     = 1, /* double only: hi/lo BF16 split (2 digits per row / column scale),
                                   3 products, FP32 accumulate, FP64 combine                          */
    MISTRA_PREC_OZAKI_BF16 = 2, /* double only: Ozaki BF16 digit slicing (5 digits, 15 products; env
                                   MISTRA_OZAKI_SLICES = 2..7), exact FP32 chunks, FP64 sum           */
    MISTRA_PREC_FP8_BLOCK  = 3, /* double only: FP8 E4M3 with one scale per row (column) and 128-long
                                   K block, FP32 accumulate within a block, FP64 across blocks        */
    MISTRA_PREC_OZAKI_INT8 = 4  /* double only: Ozaki INT8 digit slicing (6 digits of 7 bits, 21
                                   products; env MISTRA_OZAKI_I8_SLICES = 2..6), exact INT32
                                   accumulation, FP64 sum                                             */
};

/* Flags for mistra_tensor_mult. */
enum
{
    MISTRA_FLAG_REF_FRINGE = 1u << 0, /* drop the rows/cols/k that the 2^L split truncates, exactly as
                                         src/matrix/stra_tensor_view.hpp:246-247 does (default: computed) */
    MISTRA_FLAG_SYNC       = 1u << 1, /* synchronize the stream before returning                          */
    MISTRA_FLAG_GRAPH      = 1u << 2, /* mistra_batch_mult: capture the batch's launches in a HIP graph on
                                         the first call and replay it on later calls with the same entries
                                         (same pointers, shapes, scalars); needs a non-null stream        */
    MISTRA_FLAG_FUSED_SUMS = 1u << 3, /* ABC: form every operand sum inside the leaf kernel (no
                                         workspace), never staging the outer level's sums                */
    MISTRA_FLAG_OUTER_STAGED = 1u << 4, /* ABC (levels 1, 2): stage the outer level's 7 operand sums of
                                         A and of B in HBM workspace (7/4 of A's and of B's size) and form
                                         only the inner sums (2 terms at level 2, none at level 1) in the
                                         leaf kernel, whenever the shape allows it (one batch entry).   */
    MISTRA_FLAG_ALL_STAGED = 1u << 5,  /* ABC: stage the sums of every level (at level 2 the 49 sums of A
                                         and of B: 49/16 of A's and of B's size); each leaf product then
                                         reads one dense operand pair.  Without any of the three flags
                                         the cost model decides.                                         */
    MISTRA_FLAG_BATCH_GROUPED = 1u << 6, /* mistra_batch_mult: one grouped launch per (Strassen stage round,
                                         kernel variant) instead of the single persistent launch        */
    MISTRA_FLAG_STAGE_PERSIST = 1u << 7, /* ABC with every sum staged (level 1 staged, level 2 all
                                         staged): all products in one persistent launch, whose tiles
                                         wait only for earlier products updating the same C tile,
                                         instead of one leaf launch per stage (products with disjoint C
                                         targets) in stream order                                        */
    MISTRA_FLAG_ATOMIC_C = 1u << 8      /* ABC with every sum staged, FP64: all products in one launch,
                                         each adding alpha*coeff*M_p into its C sub-blocks with hardware
                                         FP64 atomic adds (C scaled by beta first).  The adds into one
                                         C element come in no fixed order: results may differ from run
                                         to run in the last bits.  Ignored where it does not apply.      */
};

/* Error codes (0 = success). */
enum
{
    MISTRA_OK              = 0,
    MISTRA_ERR_ARG         = -1, /* invalid argument (types, shapes, labels)           */
    MISTRA_ERR_UNSUPPORTED = -2, /* valid but unsupported combination (e.g. complex)   */
    MISTRA_ERR_HIP         = -3, /* a HIP runtime call failed                          */
    MISTRA_ERR_NOMEM       = -4  /* device workspace allocation failed                 */
};

/* Explicit-everything contraction: level L in {0,1,2}, impl as above, precision as above,
 * stream is a hipStream_t (NULL = the per-thread default stream). Returns 0 or a negative code. */
int mistra_tensor_mult(const tblis_tensor* A, const label_type* idx_A,
                       const tblis_tensor* B, const label_type* idx_B,
                             tblis_tensor* C, const label_type* idx_C,
                       int level, int impl, int precision, unsigned flags, void* stream);

/* mistra_tensor_mult that allocates nothing: no counterpart in the reference; it is for callers that
 * cannot allocate device memory (code between allocations of a tight serving loop, harnesses that forbid
 * hipMalloc).  The plan's tables and the staged workspaces of its launches live in a fixed 64 MiB device
 * block per GPU, held from the first call until the process ends; no hipMalloc, hipFree or pool
 * allocation happens in the call.  Same arguments and results as mistra_tensor_mult, restricted to:
 * MISTRA_PREC_NATIVE, level 0..2, impl 0..4 (no autotuning), device tensors, pure contractions.  ABC's
 * staged sums fall back to the fused form when they do not fit; AB and Naive return MISTRA_ERR_NOMEM
 * when their staged products (AB: nprod * ms * ns elements; Naive: nprod * (ms*ks + ks*ns + ms*ns))
 * exceed what the block leaves beside the tables.  Calls that overlap on one GPU (other streams or
 * threads) race on the block: serialize them.  Not capturable into a HIP graph.  Returns 0 or a negative
 * code (message in mistra_last_error). */
int mistra_tensor_mult_noalloc(const tblis_tensor* A, const label_type* idx_A,
                               const tblis_tensor* B, const label_type* idx_B,
                                     tblis_tensor* C, const label_type* idx_C,
                               int level, int impl, int precision, unsigned flags, void* stream);

/* Thread-local description of the last error of this thread ("" if none). */
const char* mistra_last_error(void);

/* Process-global impl used by the drop-in symbols (same variable as tblis::internal::impl). */
int mistra_set_impl(int impl);
int mistra_get_impl(void);

/* Release the per-device workspace arena and the plan cache. */
int mistra_release_workspace(void);

/* Kernel timing hook: when enabled, the library records HIP events around the launches of the
 * dominant (leaf-product) kernel on the stream it launches on; read back the summed time in ms,
 * the number of launches and the algorithmic flops of those launches, then reset. */
int mistra_kernel_timing(int enable);
int mistra_kernel_timing_read(double* total_ms, long long* launches, double* flops, int reset);

/* Synthetic data: fill n elements (type TYPE_FLOAT or TYPE_DOUBLE) at device pointer ptr with
 * uniform [-1,1) values from splitmix64(seed*2^32 + i); if scale_exp_max > 0, multiply each by
 * 2^u with u uniform-integer in [-scale_exp_max, scale_exp_max] (a second splitmix64 stream). */
int mistra_fill_uniform(int type, void* ptr, long long n, unsigned long long seed,
                        int scale_exp_max, void* stream);

/* Plain device copy of nbytes (a calibration kernel of known traffic for counter-based profiling). */
int mistra_copy_bytes(void* dst, const void* src, long long nbytes, void* stream);

/* Relative Frobenius distance ||X - Y||_F / ||Y||_F between two device tensors of the same
 * dense element count n (type TYPE_FLOAT or TYPE_DOUBLE; reduction in FP64 on the device). */
int mistra_rel_error(int type, const void* X, const void* Y, long long n, double* out, void* stream);

/* ---- Matrix Strassen (src/stra_iface/3m/stra_mult.h:17-27) ------------------------------------ */
/* C = alpha*A.B + beta*C for strided matrices (alpha = scalar(A)*scalar(B), beta = scalar(C)), one
 * Strassen level: ABC, Naive and AB respectively. Synchronous; errors via mistra_last_error(). */
void stra_tblis_matrix_mult(const tblis_comm* comm, const tblis_config* cfg,
                            const tblis_matrix* A, const tblis_matrix* B, tblis_matrix* C);
void stra_tblis_matrix_mult_naive(const tblis_comm* comm, const tblis_config* cfg,
                                  const tblis_matrix* A, const tblis_matrix* B, tblis_matrix* C);
void stra_tblis_matrix_mult_ab(const tblis_comm* comm, const tblis_config* cfg,
                               const tblis_matrix* A, const tblis_matrix* B, tblis_matrix* C);

/* ---- Batching: many independent contractions in one persistent launch ------------------------ */
typedef struct mistra_contraction
{
    const tblis_tensor* A; const label_type* idx_A;
    const tblis_tensor* B; const label_type* idx_B;
          tblis_tensor* C; const label_type* idx_C;
    int level;   /* -1 = autotuner's choice */
    int impl;    /* -1 = autotuner's choice (only ABC is available inside a batch) */
} mistra_contraction;

/* Run count independent contractions in one persistent launch: a device job queue first packs every
 * leaf product's operand sums (one term for a classical product) into dense K-fast panels, then runs
 * the GEMM tiles of every product, in Strassen stage order per contraction (device counters with
 * bounded waits), with the multi-C epilogue on each contraction's C.  The packed panels and tables
 * are built on the first call with a given set of entries (same pointers, shapes, strides, scalars,
 * level and impl) and kept for later calls (a few batches are cached; mistra_release_workspace drops
 * them).  MISTRA_FLAG_GRAPH records the launch in a HIP graph and replays it.  MISTRA_FLAG_SYNC also
 * reports waits that gave up (MISTRA_ERR_HIP).  MISTRA_FLAG_BATCH_GROUPED selects the older form: one
 * grouped launch per (stage round, kernel variant), ABC in its fused form.  Entries must not write
 * overlapping C memory.  level/impl -1 take the autotuner's choice; entries that are not
 * device-resident ABC contractions of the batch's element type without batch (ABC-class) indices
 * (host data, general index classes, AB/Naive) run as individual calls.  Asynchronous on stream unless
 * MISTRA_FLAG_SYNC.  No counterpart in the reference (its harness loops over contractions). */
int mistra_batch_mult(int count, const mistra_contraction* entries, unsigned flags, void* stream);

/* ---- Adjoints of C = alpha*A.B (+beta*C) ------------------------------------------------------ */
/* dA[idx_A] = alpha * sum dC[idx_C] . B[idx_B] + beta_A * dA  and
 * dB[idx_B] = alpha * sum A[idx_A] . dC[idx_C] + beta_B * dB, with alpha = scalar(A)*scalar(B)*scalar(dC)
 * (the forward alpha times a scale of dC) and beta_A, beta_B the scalars of dA, dB (0 overwrites).
 * Each is computed as a Strassen contraction of the given level/impl/precision, asynchronously on
 * stream; the scalars of dA and dB are reset to 1 as for any C. Either gradient pointer may be NULL.
 * No counterpart in the reference (its library has no adjoint entry point). */
int mistra_tensor_mult_backward(const tblis_tensor* A, const label_type* idx_A,
                                const tblis_tensor* B, const label_type* idx_B,
                                const tblis_tensor* dC, const label_type* idx_C,
                                tblis_tensor* dA, tblis_tensor* dB,
                                int level, int impl, int precision, void* stream);

/* ---- Autotuner --------------------------------------------------------------------------------- */
/* Choose level, impl and tile shape for a contraction; if measure != 0 time the candidates on the
 * device (on scratch copies of C), else use the analytic model. Results are cached per shape. */
int mistra_autotune(const tblis_tensor* A, const label_type* idx_A,
                    const tblis_tensor* B, const label_type* idx_B,
                    const tblis_tensor* C, const label_type* idx_C,
                    int measure, int* level_out, int* impl_out, int* tile_out, void* stream);

#ifdef __cplusplus
} /* extern "C" */
} /* namespace tblis */
#endif

#endif /* TBLIS_STRASSEN_AMD_H */
