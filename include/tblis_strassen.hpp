// tblis_strassen.hpp -- C++ surface of the MI355X Strassen tensor contraction library.
//
// Mirrors the templates a flame/tblis-strassen user calls:
//   tblis::stra_mult<T>(alpha, A, idx_A, B, idx_B, beta, C, idx_C)         src/stra_iface/3t/stra_mult.h:26-62
//   tblis::stra_mult_2level<T>(...)                                        src/stra_iface/3t/stra_mult_2level.h:26-62
//   tblis::mult<T>(...)                                                    src/iface/3t/mult.h:26-62
//   tblis::internal::impl (process-global impl selector)                  src/internal/3t/mult.hpp:14, mult.cxx:26
// Views: any type with data(), dimension(), lengths() and strides() works (the reference's MArray
// varray views included); tblis::tensor_view / const_tensor_view below are minimal non-owning views
// with column-major defaults (src/util/basic_types.h:21).
#pragma once

#include <cstddef>
#include <initializer_list>
#include <stdexcept>
#include <string>
#include <vector>

#include "tblis_strassen.h"

namespace tblis {

namespace internal {
// src/internal/3t/mult.hpp:14
enum impl_t
{
    BLIS_BASED,
    BLAS_BASED,
    REFERENCE,
    STRA_NAIVE,
    STRA_AB
};
extern impl_t impl;
}  // namespace internal

struct single_t
{
    constexpr single_t() {}
};
constexpr single_t single;

template <typename T>
struct type_tag;
template <>
struct type_tag<float>
{
    static constexpr type_t value = TYPE_FLOAT;
};
template <>
struct type_tag<double>
{
    static constexpr type_t value = TYPE_DOUBLE;
};

// Minimal non-owning strided view (column-major by default).
template <typename T>
class basic_view
{
public:
    basic_view() = default;
    basic_view(std::vector<len_type> len, T* data, std::vector<stride_type> stride)
        : data_(data), len_(std::move(len)), stride_(std::move(stride))
    {
        if (stride_.size() != len_.size()) throw std::invalid_argument("tblis view: lengths/strides mismatch");
    }
    basic_view(std::vector<len_type> len, T* data) : data_(data), len_(std::move(len))
    {
        stride_type s = 1;
        for (len_type l : len_)
        {
            stride_.push_back(s);
            s *= l;
        }
    }
    T* data() const { return data_; }
    unsigned dimension() const { return unsigned(len_.size()); }
    const std::vector<len_type>& lengths() const { return len_; }
    const std::vector<stride_type>& strides() const { return stride_; }
    len_type length(unsigned d) const { return len_[d]; }
    stride_type stride(unsigned d) const { return stride_[d]; }

private:
    T* data_ = nullptr;
    std::vector<len_type> len_;
    std::vector<stride_type> stride_;
};

template <typename T>
using tensor_view = basic_view<T>;
template <typename T>
using const_tensor_view = basic_view<const T>;

namespace detail {

template <typename T, typename View>
tblis_tensor make_tensor(T alpha, const View& v)
{
    tblis_tensor t;
    t.type = type_tag<T>::value;
    t.conj = 0;
    for (auto& c : t.scalar) c = 0;
    *reinterpret_cast<T*>(t.scalar) = alpha;
    t.data = const_cast<void*>(static_cast<const void*>(v.data()));
    t.ndim = v.dimension();
    t.len = const_cast<len_type*>(v.lengths().data());
    t.stride = const_cast<stride_type*>(v.strides().data());
    return t;
}

}  // namespace detail

// src/stra_iface/3t/stra_mult.h:26-36
template <typename T, typename ViewA, typename ViewB, typename ViewC>
void stra_mult(T alpha, const ViewA& A, const label_type* idx_A, const ViewB& B, const label_type* idx_B, T beta,
               const ViewC& C, const label_type* idx_C)
{
    tblis_tensor A_s = detail::make_tensor(alpha, A);
    tblis_tensor B_s = detail::make_tensor(T(1), B);
    tblis_tensor C_s = detail::make_tensor(beta, C);
    stra_tblis_tensor_mult(nullptr, nullptr, &A_s, idx_A, &B_s, idx_B, &C_s, idx_C);
}

template <typename T, typename ViewA, typename ViewB, typename ViewC>
void stra_mult(single_t, T alpha, const ViewA& A, const label_type* idx_A, const ViewB& B, const label_type* idx_B,
               T beta, const ViewC& C, const label_type* idx_C)
{
    tblis_tensor A_s = detail::make_tensor(alpha, A);
    tblis_tensor B_s = detail::make_tensor(T(1), B);
    tblis_tensor C_s = detail::make_tensor(beta, C);
    stra_tblis_tensor_mult(tblis_single, nullptr, &A_s, idx_A, &B_s, idx_B, &C_s, idx_C);
}

// src/stra_iface/3t/stra_mult_2level.h:26-36
template <typename T, typename ViewA, typename ViewB, typename ViewC>
void stra_mult_2level(T alpha, const ViewA& A, const label_type* idx_A, const ViewB& B, const label_type* idx_B,
                      T beta, const ViewC& C, const label_type* idx_C)
{
    tblis_tensor A_s = detail::make_tensor(alpha, A);
    tblis_tensor B_s = detail::make_tensor(T(1), B);
    tblis_tensor C_s = detail::make_tensor(beta, C);
    stra_tblis_tensor_mult_2level(nullptr, nullptr, &A_s, idx_A, &B_s, idx_B, &C_s, idx_C);
}

template <typename T, typename ViewA, typename ViewB, typename ViewC>
void stra_mult_2level(single_t, T alpha, const ViewA& A, const label_type* idx_A, const ViewB& B,
                      const label_type* idx_B, T beta, const ViewC& C, const label_type* idx_C)
{
    tblis_tensor A_s = detail::make_tensor(alpha, A);
    tblis_tensor B_s = detail::make_tensor(T(1), B);
    tblis_tensor C_s = detail::make_tensor(beta, C);
    stra_tblis_tensor_mult_2level(tblis_single, nullptr, &A_s, idx_A, &B_s, idx_B, &C_s, idx_C);
}

// src/iface/3t/mult.h:26-36
template <typename T, typename ViewA, typename ViewB, typename ViewC>
void mult(T alpha, const ViewA& A, const label_type* idx_A, const ViewB& B, const label_type* idx_B, T beta,
          const ViewC& C, const label_type* idx_C)
{
    tblis_tensor A_s = detail::make_tensor(alpha, A);
    tblis_tensor B_s = detail::make_tensor(T(1), B);
    tblis_tensor C_s = detail::make_tensor(beta, C);
    tblis_tensor_mult(nullptr, nullptr, &A_s, idx_A, &B_s, idx_B, &C_s, idx_C);
}

// Extension: explicit level / impl / precision / stream, throwing on error.
template <typename T, typename ViewA, typename ViewB, typename ViewC>
void stra_mult_ex(T alpha, const ViewA& A, const label_type* idx_A, const ViewB& B, const label_type* idx_B, T beta,
                  const ViewC& C, const label_type* idx_C, int level, int impl = MISTRA_IMPL_ABC,
                  int precision = MISTRA_PREC_NATIVE, void* stream = nullptr, unsigned flags = 0)
{
    tblis_tensor A_s = detail::make_tensor(alpha, A);
    tblis_tensor B_s = detail::make_tensor(T(1), B);
    tblis_tensor C_s = detail::make_tensor(beta, C);
    int rc = mistra_tensor_mult(&A_s, idx_A, &B_s, idx_B, &C_s, idx_C, level, impl, precision, flags, stream);
    if (rc != MISTRA_OK) throw std::runtime_error(std::string("mistra_tensor_mult: ") + mistra_last_error());
}

// Matrix Strassen, src/stra_iface/3m/stra_mult.h:29-100: any view with data(), length(d), stride(d).
namespace detail {
template <typename T, typename View>
tblis_matrix make_matrix(T alpha, const View& v)
{
    tblis_matrix t;
    t.type = type_tag<T>::value;
    t.conj = 0;
    for (auto& c : t.scalar) c = 0;
    *reinterpret_cast<T*>(t.scalar) = alpha;
    t.data = const_cast<void*>(static_cast<const void*>(v.data()));
    t.m = v.length(0);
    t.n = v.length(1);
    t.rs = v.stride(0);
    t.cs = v.stride(1);
    return t;
}
template <typename T, typename VA, typename VB, typename VC>
void matrix_call(void (*fn)(const tblis_comm*, const tblis_config*, const tblis_matrix*, const tblis_matrix*,
                            tblis_matrix*),
                 T alpha, const VA& A, const VB& B, T beta, const VC& C)
{
    tblis_matrix A_s = make_matrix(alpha, A), B_s = make_matrix(T(1), B), C_s = make_matrix(beta, C);
    fn(nullptr, nullptr, &A_s, &B_s, &C_s);
    const char* err = mistra_last_error();
    if (err && *err) throw std::runtime_error(err);
}
}  // namespace detail

template <typename T, typename VA, typename VB, typename VC>
void stra_mult(T alpha, const VA& A, const VB& B, T beta, const VC& C)
{
    detail::matrix_call(stra_tblis_matrix_mult, alpha, A, B, beta, C);
}
template <typename T, typename VA, typename VB, typename VC>
void stra_mult_naive(T alpha, const VA& A, const VB& B, T beta, const VC& C)
{
    detail::matrix_call(stra_tblis_matrix_mult_naive, alpha, A, B, beta, C);
}
template <typename T, typename VA, typename VB, typename VC>
void stra_mult_ab(T alpha, const VA& A, const VB& B, T beta, const VC& C)
{
    detail::matrix_call(stra_tblis_matrix_mult_ab, alpha, A, B, beta, C);
}

// Extension: adjoints dA = alpha * dC . B^T and dB = alpha * A^T . dC as Strassen contractions.
template <typename T, typename ViewA, typename ViewB, typename ViewC, typename ViewGA, typename ViewGB>
void stra_mult_backward(T alpha, const ViewA& A, const label_type* idx_A, const ViewB& B, const label_type* idx_B,
                        const ViewC& dC, const label_type* idx_C, const ViewGA* dA, const ViewGB* dB, int level,
                        int impl = MISTRA_IMPL_ABC, int precision = MISTRA_PREC_NATIVE, void* stream = nullptr)
{
    tblis_tensor A_s = detail::make_tensor(T(1), A);
    tblis_tensor B_s = detail::make_tensor(T(1), B);
    tblis_tensor C_s = detail::make_tensor(alpha, dC);
    tblis_tensor gA, gB;
    if (dA) gA = detail::make_tensor(T(0), *dA);
    if (dB) gB = detail::make_tensor(T(0), *dB);
    int rc = mistra_tensor_mult_backward(&A_s, idx_A, &B_s, idx_B, &C_s, idx_C, dA ? &gA : nullptr,
                                         dB ? &gB : nullptr, level, impl, precision, stream);
    if (rc != MISTRA_OK) throw std::runtime_error(std::string("mistra_tensor_mult_backward: ") + mistra_last_error());
}

}  // namespace tblis
