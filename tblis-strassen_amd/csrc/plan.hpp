// Host-side index layer and Strassen plan of the MI355X tensor-contraction library.
//
// Restates, in this library's own code, the index processing of flame/tblis-strassen:
//   diagonal / fold / sort_by_stride   src/util/tensor.hpp:305-342, :270-303, :37-78
//   7-way index grouping               src/stra_iface/3t/stra_mult.cxx:35-180
//   per-group stride ordering          src/stra_internal/3t/stra_mult.cxx:336-356
//   scatter vectors                    src/matrix/stra_tensor_view.hpp:427-455 (fill_scatter)
//   block-scatter ("run") strides      src/matrix/stra_tensor_view.hpp:457-513 (fill_block_scatter)
//   sub-block partition                src/matrix/stra_tensor_view.hpp:198-257 (divisor, subid offsets)
//   1-level Strassen table             src/stra_internal/3t/strassen:1-14, stra_mult.cxx:426-591
//   2-level table (Kronecker square)   src/stra_internal/3t/stra_mult_2level.cxx:341-1458
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "tblis_strassen.h"

namespace mistra {

using tblis::len_type;
using tblis::stride_type;
using tblis::label_type;
// the C-ABI enums (MISTRA_*, TYPE_*) live in namespace tblis when compiled as C++
using namespace ::tblis;

// One index group after grouping and folding: lengths plus the strides in up to three operands.
struct IndexGroup
{
    std::vector<label_type> idx;
    std::vector<len_type> len;
    std::vector<stride_type> sA, sB, sC;
};

// Result of the index analysis of one contraction (labels merged, lengths-1 dropped, groups folded).
struct Grouping
{
    IndexGroup AB, AC, BC, ABC, A_only, B_only, C_only;
    // true when Strassen applies: no A-only/B-only/C-only indices and (AB or ABC) empty
    // (src/stra_internal/3t/stra_mult.cxx:1728-1731).  Batch (ABC) indices with a non-empty AB group
    // go to the general path, which this library still runs through the same kernels.
    bool pure() const
    {
        return A_only.len.empty() && B_only.len.empty() && C_only.len.empty();
    }
};

// diagonal(): sort dims by label, merge repeated labels by summing strides, drop length-1 dims.
void diagonal(unsigned ndim, const len_type* len, const stride_type* stride, const label_type* idx,
              std::vector<len_type>& len_out, std::vector<stride_type>& stride_out,
              std::vector<label_type>& idx_out);

// Analyse a contraction: returns false (with a message) on a label/length mismatch.
bool group_indices(const tblis::tblis_tensor& A, const label_type* idx_A,
                   const tblis::tblis_tensor& B, const label_type* idx_B,
                   const tblis::tblis_tensor& C, const label_type* idx_C,
                   Grouping& g, std::string& err);

// fold(): merge indices that are contiguous in every stride vector (ordered by the first).
void fold(IndexGroup& g, int nstrides);

// Stable lexicographic ordering by (primary, secondary) stride, as sort_by_stride.
std::vector<unsigned> sort_by_stride(const std::vector<stride_type>& s1, const std::vector<stride_type>& s2);

// Scatter vector of a flattened index group: pos -> sum_d digit_d(pos) * stride_d, first index fastest.
std::vector<int64_t> scatter(const std::vector<len_type>& len, const std::vector<stride_type>& stride);

// Run stride of a scatter vector for runs of length `run`: out[p] = s if scat[p+j] = scat[p] + j*s for
// every j < run with p+j < n, else 0 (a block-scatter entry that any position, aligned or not, can use).
std::vector<int64_t> run_strides(const std::vector<int64_t>& scat, int run);

// ---- Strassen tables -------------------------------------------------------------------------------

struct SubTerm
{
    int subid;     // sub-block id: row block * divisor + col block
    double coeff;  // +1 / -1 (products of +-1 at two levels)
};

struct Product
{
    std::vector<SubTerm> A, B, C;
};

// 1-level: the 7 products in the reference order (M0..M6), sub-block ids 0..3 (row-major 2x2).
const std::vector<Product>& strassen_products_1level();
// 2-level: the 49 products, product 7*p_outer + p_inner, sub-block id 4*(2 s1 + s2) + (2 t1 + t2),
// terms ordered outer-major, coefficient = outer * inner.
const std::vector<Product>& strassen_products_2level();
// Level 0: a single product with identity terms.
const std::vector<Product>& products_for_level(int level);

}  // namespace mistra
