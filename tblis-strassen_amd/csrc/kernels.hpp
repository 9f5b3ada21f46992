// Descriptors shared by the host driver and the gfx950 kernels.
//
// A contraction is flattened to C(m,n) += A(m,k) B(k,n) over *scatter vectors*: element (r,c) of an
// operand lives at data + rs[r] + cs[c].  This is the device form of the reference's block-scatter
// matrices (src/matrix/stra_block_scatter_matrix.hpp:9-306, filled per call by
// src/stra_nodes/stra_matrify.hpp:80-412); here the vectors are built once per contraction and stay
// in HBM.  A Strassen sub-block is a (row offset, col offset) into those vectors plus a coefficient,
// which is what src/matrix/stra_tensor_view.hpp:246-247 computes from a sub-block id.
#pragma once

#include <cstdint>

namespace mistra {

constexpr int kMaxTerms = 4;  // at most 4 sub-blocks per operand sum (2-level Strassen)

struct SubRef
{
    int64_t base;   // element offset added to the data pointer (per-product staging buffers)
    int64_t roff;   // offset into the row scatter vector
    int64_t coff;   // offset into the column scatter vector
    double coeff;   // coefficient of this sub-block in the sum
};

struct ProdDesc
{
    int nA, nB, nC;
    int first_mask;  // bit c set: this product is the first writer of C target c (beta applied here)
    SubRef a[kMaxTerms], b[kMaxTerms], c[kMaxTerms];
};

// Epilogue kinds of the leaf GEMM kernel.
enum EpiKind
{
    EPI_MULTI_C = 0,  // C_c = (first ? beta*C_c : C_c) + alpha*coeff_c*M for each target (ABC, classical)
    EPI_STORE = 1,    // M_p stored dense, column-major, ld = ms, product p at M + p*ms*ns (AB / Naive)
    EPI_ATOMIC_C = 2  // C_c += alpha*coeff_c*M for each target by floating-point atomic adds: the products of
                      // one launch may share targets (ABC in one launch; beta applied to C beforehand)
};

// How the leaf kernel addresses K (chosen per launch by the host).
enum KMode
{
    KMODE_GENERAL = 0,  // per-term gathered K positions (any sub-block layout)
    KMODE_AFFINE = 1,   // translate sub-blocks, K runs affine per K-tile
    KMODE_GATHER = 2,   // translate sub-blocks, K offsets gathered per lane
    KMODE_AFFINE_VEC = 3  // KMODE_AFFINE with ks a multiple of the K-tile: K-fast operands (unit-stride
                          // K) load 16 bytes per lane
};

struct GemmArgs
{
    const void* A;
    const void* B;
    void* C;
    // scatter vectors (element offsets) and K-run strides
    const int64_t* a_m;
    const int64_t* a_k;
    const int64_t* a_krun;
    const int64_t* b_k;
    const int64_t* b_krun;
    const int64_t* b_n;
    const int64_t* c_m;
    const int64_t* c_n;
    int64_t ms, ns, ks;  // leaf sub-problem extent
    const ProdDesc* prods;
    int nprod;
    int tiles_m, tiles_n;
    double alpha, beta;
    // batch (ABC) indices: entry z adds batch_a[z], batch_b[z], batch_c[z] to the data pointers
    int64_t nbatch;
    const int64_t* batch_a;
    const int64_t* batch_b;
    const int64_t* batch_c;
    // staging buffer of EPI_STORE
    void* M;
    // a few zeroed bytes in device memory, read by lanes whose position lies outside the leaf
    const void* zero;
    // fast K mode (chosen by the host per launch): every term's K-tiles are affine with these strides
    // and every element offset fits in 31 bits
    int64_t ka_stride, kb_stride;
    // direct-to-LDS body: 0 = C targets are shifts (ProdDesc base) of the first C sub-block; 1 = they are
    // addressed through (roff, coff) into c_m / c_n (sub-blocks that are not translates of each other)
    int c_scatter;
    // multi-C epilogue: C is read and written with device-coherent (agent-scope) accesses, because
    // workgroups on other XCDs read or write the same C during the same launch (the rounds of one
    // contraction in the persistent batch launch); 0 = ordinary cached accesses
    int c_coherent;
};

}  // namespace mistra
