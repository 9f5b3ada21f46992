// A C++ caller of the template surface (include/tblis_strassen.hpp), written the way a user of the
// reference's templates calls them: tblis::stra_mult<T> (src/stra_iface/3t/stra_mult.h:26-62),
// tblis::stra_mult_2level<T> (src/stra_iface/3t/stra_mult_2level.h:26-62) and tblis::mult<T>
// (src/iface/3t/mult.h:26-62), with the impl chosen through the process-global tblis::internal::impl
// (src/internal/3t/mult.cxx:26).  Inputs are small integers, so every level and impl must reproduce the
// brute-force contraction below bit for bit.  Host views (staged by the library) and device views
// (stra_mult_ex, the matrix overloads) both run.  Prints one line per check and "ALL OK" at the end;
// exit status 0 only if every check passed.
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstdlib>
#include <map>
#include <string>
#include <vector>

#include "tblis_strassen.hpp"

namespace {

int failures = 0;

template <typename T>
std::vector<T> small_ints(size_t n, unsigned seed)
{
    std::vector<T> v(n);
    unsigned s = seed * 2654435761u + 12345u;
    for (auto& x : v)
    {
        s = s * 1664525u + 1013904223u;
        x = T(int((s >> 16) % 7u) - 3);
    }
    return v;
}

struct Case
{
    std::string ia, ib, ic;
    std::map<char, long> len;
    std::vector<long> shape(const std::string& idx) const
    {
        std::vector<long> s;
        for (char c : idx) s.push_back(len.at(c));
        return s;
    }
    long volume(const std::string& idx) const
    {
        long v = 1;
        for (char c : idx) v *= len.at(c);
        return v;
    }
};

// column-major offset of a label assignment
long offset(const std::string& idx, const Case& cs, const std::map<char, long>& at)
{
    long off = 0, st = 1;
    for (char c : idx)
    {
        off += at.at(c) * st;
        st *= cs.len.at(c);
    }
    return off;
}

// C = alpha * sum A*B + beta * C over every label assignment (labels of C and the summed ones)
template <typename T>
std::vector<T> brute(const Case& cs, const std::vector<T>& A, const std::vector<T>& B, std::vector<T> C, T alpha,
                     T beta)
{
    std::vector<char> labels;
    for (auto& kv : cs.len) labels.push_back(kv.first);
    std::vector<T> acc(C.size(), T(0));
    std::map<char, long> at;
    for (char c : labels) at[c] = 0;
    while (true)
    {
        acc[size_t(offset(cs.ic, cs, at))] += A[size_t(offset(cs.ia, cs, at))] * B[size_t(offset(cs.ib, cs, at))];
        size_t d = 0;
        for (; d < labels.size(); d++)
        {
            if (++at[labels[d]] < cs.len.at(labels[d])) break;
            at[labels[d]] = 0;
        }
        if (d == labels.size()) break;
    }
    for (size_t i = 0; i < C.size(); i++) C[i] = alpha * acc[i] + beta * C[i];
    return C;
}

template <typename T>
void check(const char* what, const std::vector<T>& got, const std::vector<T>& want)
{
    size_t bad = 0;
    for (size_t i = 0; i < got.size(); i++)
        if (got[i] != want[i]) bad++;
    std::printf("%-44s %s (%zu of %zu differ)\n", what, bad ? "FAIL" : "ok", bad, got.size());
    const char* err = tblis::mistra_last_error();
    if (err && *err) std::printf("    last error: %s\n", err);
    if (bad) failures++;
}

template <typename T>
void host_templates(const Case& cs, const char* tname)
{
    const auto a = small_ints<T>(size_t(cs.volume(cs.ia)), 1), b = small_ints<T>(size_t(cs.volume(cs.ib)), 2),
               c0 = small_ints<T>(size_t(cs.volume(cs.ic)), 3);
    const T alpha = T(2), beta = T(-1);
    const auto want = brute(cs, a, b, c0, alpha, beta);
    tblis::const_tensor_view<T> A(cs.shape(cs.ia), a.data()), B(cs.shape(cs.ib), b.data());
    const struct
    {
        tblis::internal::impl_t impl;
        const char* name;
    } impls[] = {{tblis::internal::BLIS_BASED, "BLIS_BASED"},
                 {tblis::internal::STRA_AB, "STRA_AB"},
                 {tblis::internal::STRA_NAIVE, "STRA_NAIVE"}};
    for (const auto& im : impls)
    {
        tblis::internal::impl = im.impl;
        char what[128];
        auto c = c0;
        tblis::tensor_view<T> C(cs.shape(cs.ic), c.data());
        tblis::stra_mult<T>(alpha, A, cs.ia.c_str(), B, cs.ib.c_str(), beta, C, cs.ic.c_str());
        std::snprintf(what, sizeof(what), "stra_mult<%s> %s", tname, im.name);
        check(what, c, want);
        c = c0;
        tblis::stra_mult_2level<T>(alpha, A, cs.ia.c_str(), B, cs.ib.c_str(), beta, C, cs.ic.c_str());
        std::snprintf(what, sizeof(what), "stra_mult_2level<%s> %s", tname, im.name);
        check(what, c, want);
        c = c0;
        tblis::stra_mult<T>(tblis::single, alpha, A, cs.ia.c_str(), B, cs.ib.c_str(), beta, C, cs.ic.c_str());
        std::snprintf(what, sizeof(what), "stra_mult<%s>(single) %s", tname, im.name);
        check(what, c, want);
        c = c0;
        tblis::mult<T>(alpha, A, cs.ia.c_str(), B, cs.ib.c_str(), beta, C, cs.ic.c_str());
        std::snprintf(what, sizeof(what), "mult<%s> %s", tname, im.name);
        check(what, c, want);
    }
    tblis::internal::impl = tblis::internal::BLIS_BASED;
}

template <typename T>
T* to_device(const std::vector<T>& v)
{
    void* p = nullptr;
    if (hipMalloc(&p, v.size() * sizeof(T)) != hipSuccess ||
        hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
    {
        std::printf("device allocation failed\n");
        std::exit(2);
    }
    return static_cast<T*>(p);
}

template <typename T>
std::vector<T> from_device(const T* p, size_t n)
{
    std::vector<T> v(n);
    if (hipMemcpy(v.data(), p, n * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess) std::exit(2);
    return v;
}

void device_views(const Case& cs)
{
    const auto a = small_ints<double>(size_t(cs.volume(cs.ia)), 4), b = small_ints<double>(size_t(cs.volume(cs.ib)), 5),
               c0 = small_ints<double>(size_t(cs.volume(cs.ic)), 6);
    const auto want = brute(cs, a, b, c0, 3.0, 0.5);
    double *dA = to_device(a), *dB = to_device(b);
    tblis::const_tensor_view<double> A(cs.shape(cs.ia), dA), B(cs.shape(cs.ib), dB);
    const int impls[] = {tblis::MISTRA_IMPL_ABC, tblis::MISTRA_IMPL_AB, tblis::MISTRA_IMPL_NAIVE};
    for (int level = 0; level <= 2; level++)
        for (int impl : impls)
        {
            double* dC = to_device(c0);
            tblis::tensor_view<double> C(cs.shape(cs.ic), dC);
            tblis::stra_mult_ex(3.0, A, cs.ia.c_str(), B, cs.ib.c_str(), 0.5, C, cs.ic.c_str(), level, impl,
                                tblis::MISTRA_PREC_NATIVE, nullptr, tblis::MISTRA_FLAG_SYNC);
            char what[96];
            std::snprintf(what, sizeof(what), "stra_mult_ex<double> device L%d impl %d", level, impl);
            check(what, from_device(dC, c0.size()), want);
            (void)hipFree(dC);
        }
    (void)hipFree(dA);
    (void)hipFree(dB);
}

void matrix_overloads()
{
    // C (m x n) = 2 A (m x k) B (k x n) - C, A row-major, B and C column-major, device memory
    const long m = 70, n = 54, k = 62;
    const auto a = small_ints<double>(size_t(m * k), 7), b = small_ints<double>(size_t(k * n), 8),
               c0 = small_ints<double>(size_t(m * n), 9);
    std::vector<double> want(c0.size());
    for (long i = 0; i < m; i++)
        for (long j = 0; j < n; j++)
        {
            double s = 0;
            for (long p = 0; p < k; p++) s += a[size_t(i * k + p)] * b[size_t(p + j * k)];
            want[size_t(i + j * m)] = 2.0 * s - c0[size_t(i + j * m)];
        }
    double *dA = to_device(a), *dB = to_device(b);
    tblis::const_tensor_view<double> A({m, k}, dA, {k, 1}), B({k, n}, dB, {1, k});
    for (int which = 0; which < 3; which++)
    {
        double* dC = to_device(c0);
        tblis::tensor_view<double> C({m, n}, dC, {1, m});
        if (which == 0) tblis::stra_mult(2.0, A, B, -1.0, C);
        if (which == 1) tblis::stra_mult_naive(2.0, A, B, -1.0, C);
        if (which == 2) tblis::stra_mult_ab(2.0, A, B, -1.0, C);
        const char* names[] = {"matrix stra_mult", "matrix stra_mult_naive", "matrix stra_mult_ab"};
        check(names[which], from_device(dC, c0.size()), want);
        (void)hipFree(dC);
    }
    (void)hipFree(dA);
    (void)hipFree(dB);
}

}  // namespace

int main()
{
    // C[q,j,i] = sum_{k,l} A[k,i,q,l] B[l,j,k]: m = {i, q}, n = {j}, k = {k, l}, odd lengths (fringes)
    Case cs{"kiql", "ljk", "qji", {{'i', 10}, {'q', 7}, {'j', 13}, {'k', 9}, {'l', 6}}};
    host_templates<double>(cs, "double");
    host_templates<float>(cs, "float");
    device_views(cs);
    matrix_overloads();
    std::printf(failures ? "FAILED: %d checks\n" : "ALL OK\n", failures);
    return failures ? 1 : 0;
}
