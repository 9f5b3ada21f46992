/*
 * tblis_strassen_debug.h -- introspection of the host plan, for tests.  No reference counterpart:
 * it exposes the index analysis that src/stra_iface/3t/stra_mult.cxx:35-180 and
 * src/stra_internal/3t/stra_mult.cxx:336-407 perform internally, without touching the device.
 */
#ifndef TBLIS_STRASSEN_AMD_DEBUG_H
#define TBLIS_STRASSEN_AMD_DEBUG_H

#include "tblis_strassen.h"

#ifdef __cplusplus
namespace tblis {
extern "C" {
#endif

/* dims[0..15] = m, n, k, nbatch, pure, transposed, ms, ns, ks, nprod (core products), nfringe,
 *               a_mfast, b_nfast, staged_depth (levels whose operand sums are staged), number of ABC
 *               launch stages, staged (1: operand sums, or at level 0 packed copies, staged in HBM).
 * scat[6*cap]  = a_m, a_k, b_k, b_n, c_m, c_n (in the plan's own orientation), -1 padded.
 * prods[cap_prods*40] per product (core then fringe): nA, nB, nC, first_mask, then 4 x (roff, coff,
 *               coeff) for A, B and C (-1, -1, 0 padded).  Returns the number of products or < 0. */
int mistra_debug_plan(const tblis_tensor* A, const label_type* idx_A,
                      const tblis_tensor* B, const label_type* idx_B,
                      const tblis_tensor* C, const label_type* idx_C,
                      int level, unsigned flags, long long* dims, long long* scat, long long cap,
                      double* prods, int cap_prods);

/* Run strides of a scatter vector (see run_strides in the plan code); out has n entries. */
int mistra_debug_run_strides(const long long* scat, long long n, int run, long long* out);

/* Combine pass of the staged products into C (AB, Naive, low precision): 0 = the one-pass form where the
 * plan allows it (every staged product read once, all C sub-blocks updated), 1 = always the per-sub-block
 * form.  Both sum each sub-block's terms in the same order, so results are identical; the switch exists
 * to test that.  Process-wide; returns the previous setting. */
int mistra_debug_set_combine_form(int per_sub_block);

/* Digit counts of the Ozaki modes for every later call of this process, overriding the environment
 * (MISTRA_OZAKI_SLICES, MISTRA_OZAKI_I8_SLICES): bf16_slices 2..7 (6 and 7 take the GEMM kernel that
 * stages through registers), int8_slices 2..6, and int8_fold_ksteps 1..1024: the INT8 GEMM folds its
 * exact INT32 sums into FP64 every that many 64-deep K-steps (by default only for K beyond 65,536).
 * 0 restores a default.  Returns 0, or MISTRA_ERR_ARG for a value out of range. */
int mistra_debug_set_lowp(int bf16_slices, int int8_slices, int int8_fold_ksteps);

/* Waits of the persistent launches (the batch launch and the single launch over a staged plan's
 * products): spin_limit polls before a wait gives up (0 = the default) and dep_bias added to every
 * wait's target (0 in normal use; a positive bias makes every wait unreachable, so each gives up after
 * spin_limit polls and the launch still ends).  Process-wide; returns 0. */
int mistra_debug_set_batch_wait(long long spin_limit, unsigned dep_bias);

/* The autotuner's candidates for a contraction shape (host-side model only, no device work): level, impl,
 * register tile and modelled milliseconds of each, fastest first, up to cap entries.  Returns the number
 * of candidates (0 for contractions that take the general path), or a negative error code. */
int mistra_debug_tune_candidates(const tblis_tensor* A, const label_type* idx_A, const tblis_tensor* B,
                                 const label_type* idx_B, const tblis_tensor* C, const label_type* idx_C, int* level,
                                 int* impl, int* tile, double* model_ms, int cap);

/* Low-precision plans of one level whose operand is K-fast: 1 (default) forms the operand's 7 sums and
 * their digits in one launch; 0 stages the sums in HBM and slices them in a second launch.  Both give the
 * same digits bit for bit; the switch exists to test that.  Process-wide; returns 0. */
int mistra_debug_set_lowp_fusion(int fused);

/* Register tile of the leaf launches of every later call that does not name one (plans built from then
 * on; cached plans keep their tile): 0 = 128x128, 1 = 64x64, 2 = 128x64 (leaf rows x columns), 3 = 64x128
 * (the rectangular tiles serve single-term launches; the plan's other launches take 128x128), -1 = the
 * plan's own choice.  For tuning sweeps and tests.  Process-wide; returns 0, or MISTRA_ERR_ARG for an
 * unknown tile. */
int mistra_debug_set_tile(int tile);

/* Builds the persistent batch launch's tables for these entries on the host alone (plans without device
 * tables, placeholder addresses; data pointers are not dereferenced) and replays its job queue in order:
 * every tile's waits must be met by the jobs before it and every product's pack count must come out
 * exact.  Returns 0, or a negative code with the first violation in msg. */
int mistra_debug_batch_check(int count, const mistra_contraction* entries, char* msg, int msg_len);

#ifdef __cplusplus
} /* extern "C" */
} /* namespace tblis */
#endif

#endif
