// C-ABI of the MI355X Strassen tensor contraction library.
//
// stra_tblis_tensor_mult / _2level / tblis_tensor_mult keep the reference's signatures and grouping
// semantics (src/stra_iface/3t/stra_mult.cxx:26-254, src/iface/3t/mult.cxx:15-146): alpha = alpha_A *
// alpha_B, beta = alpha_C, alpha == 0 scales (or zeroes) C, and on return C's scalar is 1 and conj 0.
// Differences, by design: alpha/beta hold for every impl (the reference is consistent only at
// alpha = beta = 1), every row/column/k is computed (no truncation by the 2^L split), and errors are
// reported (thread-local message, int codes on the mistra_ extensions) instead of undefined behaviour.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "driver.hpp"
#include "tblis_strassen.hpp"

namespace tblis {
namespace internal {
impl_t impl = BLIS_BASED;  // src/internal/3t/mult.cxx:26
}  // namespace internal

namespace {
const char g_single_tag = 0;
}
extern "C" const tblis_comm* const tblis_single = reinterpret_cast<const tblis_comm*>(&g_single_tag);

}  // namespace tblis

namespace mistra {

namespace {

thread_local std::string g_last_error;

int fail(const Status& s)
{
    g_last_error = s.msg;
    return s.code;
}

// ---- kernel timing hook -------------------------------------------------------------------------------

struct TimedLaunch
{
    hipEvent_t start, stop;
    double flops;
};
std::mutex g_timing_mutex;
bool g_timing = false;
std::vector<TimedLaunch> g_pending;
hipEvent_t g_cur_start = nullptr;

// ---- helpers ---------------------------------------------------------------------------------------------

bool is_device_pointer(const void* p)
{
    if (!p) return true;
    hipPointerAttribute_t attr;
    hipError_t e = hipPointerGetAttributes(&attr, p);
    if (e != hipSuccess)
    {
        (void)hipGetLastError();
        return false;
    }
    return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

// [min, max] element offsets touched by a strided tensor
void span(const tblis::tblis_tensor* t, int64_t& lo, int64_t& hi)
{
    lo = hi = 0;
    for (unsigned i = 0; i < t->ndim; i++)
    {
        if (t->len[i] <= 0) continue;
        const int64_t ext = int64_t(t->len[i] - 1) * int64_t(t->stride[i]);
        if (ext < 0) lo += ext;
        else hi += ext;
    }
}

bool empty_tensor(const tblis::tblis_tensor* t)
{
    for (unsigned i = 0; i < t->ndim; i++)
        if (t->len[i] == 0) return true;
    return false;
}

template <typename T>
T scalar_of(const tblis::tblis_tensor* t)
{
    T v;
    std::memcpy(&v, t->scalar, sizeof(T));
    return v;
}

double scalar_as_double(const tblis::tblis_tensor* t)
{
    return t->type == tblis::TYPE_DOUBLE ? scalar_of<double>(t) : double(scalar_of<float>(t));
}

}  // namespace

void set_scalar(tblis::tblis_tensor* t, double v)
{
    if (t->type == tblis::TYPE_DOUBLE)
        std::memcpy(t->scalar, &v, sizeof(v));
    else
    {
        const float f = float(v);
        std::memcpy(t->scalar, &f, sizeof(f));
    }
}

namespace {

void set_scalar_one(tblis::tblis_tensor* t)
{
    if (t->type == tblis::TYPE_DOUBLE)
    {
        const double one = 1.0;
        std::memcpy(t->scalar, &one, sizeof(one));
    }
    else if (t->type == tblis::TYPE_FLOAT)
    {
        const float one = 1.0f;
        std::memcpy(t->scalar, &one, sizeof(one));
    }
    t->conj = 0;
}

Status scale_c(const tblis::tblis_tensor* C, void* data, double beta, hipStream_t s)
{
    StridedDims d;
    std::memset(&d, 0, sizeof(d));
    if (C->ndim > 8) return Status::error(MISTRA_ERR_UNSUPPORTED, "C has more than 8 dimensions");
    d.ndim = int(C->ndim);
    for (unsigned i = 0; i < C->ndim; i++)
    {
        d.len[i] = C->len[i];
        d.stride[i] = C->stride[i];
    }
    if (beta == 1.0) return Status();
    hipError_t e = launch_scale(C->type == tblis::TYPE_DOUBLE ? 8 : 4, data, d, beta, s);
    if (e != hipSuccess) return Status::error(MISTRA_ERR_HIP, hipGetErrorString(e));
    return Status();
}

Status validate(const tblis::tblis_tensor* A, const label_type* idx_A, const tblis::tblis_tensor* B,
                const label_type* idx_B, const tblis::tblis_tensor* C, const label_type* idx_C)
{
    if (!A || !B || !C) return Status::error(MISTRA_ERR_ARG, "null tensor");
    for (const tblis::tblis_tensor* t : {A, B, C})
    {
        if (t->ndim > 0 && (!t->len || !t->stride)) return Status::error(MISTRA_ERR_ARG, "null len/stride");
        for (unsigned i = 0; i < t->ndim; i++)
            if (t->len[i] < 0) return Status::error(MISTRA_ERR_ARG, "negative length");
    }
    if ((A->ndim && !idx_A) || (B->ndim && !idx_B) || (C->ndim && !idx_C))
        return Status::error(MISTRA_ERR_ARG, "null index string");
    return Status();
}

}  // namespace

void timing_begin(hipStream_t s)
{
    std::lock_guard<std::mutex> lk(g_timing_mutex);
    if (!g_timing) return;
    if (hipEventCreate(&g_cur_start) != hipSuccess)
    {
        g_cur_start = nullptr;
        return;
    }
    (void)hipEventRecord(g_cur_start, s);
}

void timing_end(hipStream_t s, double flops)
{
    std::lock_guard<std::mutex> lk(g_timing_mutex);
    if (!g_timing || !g_cur_start) return;
    hipEvent_t stop;
    if (hipEventCreate(&stop) != hipSuccess) return;
    (void)hipEventRecord(stop, s);
    g_pending.push_back(TimedLaunch{g_cur_start, stop, flops});
    g_cur_start = nullptr;
}

Status tensor_mult(const tblis::tblis_tensor* A, const label_type* idx_A, const tblis::tblis_tensor* B,
                   const label_type* idx_B, tblis::tblis_tensor* C, const label_type* idx_C, int level, int impl,
                   int precision, unsigned flags, hipStream_t stream, bool synchronous)
{
    Status st = validate(A, idx_A, B, idx_B, C, idx_C);
    if (!st.ok()) return st;
    if (precision != MISTRA_PREC_NATIVE)
        return Status::error(MISTRA_ERR_UNSUPPORTED, "precision mode not available in this build");
    if (impl == MISTRA_IMPL_AUTO) impl = IMPL_ABC;
    if (impl < 0 || impl > 4) return Status::error(MISTRA_ERR_ARG, "impl must be in 0..4");

    const double alpha = scalar_as_double(A) * scalar_as_double(B);
    const double beta = scalar_as_double(C);

    std::shared_ptr<Plan> plan;
    if (!empty_tensor(C) && alpha != 0.0 && !empty_tensor(A) && !empty_tensor(B))
    {
        st = get_plan(A, idx_A, B, idx_B, C, idx_C, level, impl, -1, flags, plan);
        if (!st.ok()) return st;
    }
    else if (A->type != C->type || B->type != C->type)
        return Status::error(MISTRA_ERR_ARG, "A, B and C must have the same type");

    const int elem = C->type == tblis::TYPE_DOUBLE ? 8 : (C->type == tblis::TYPE_FLOAT ? 4 : 0);
    if (!elem) return Status::error(MISTRA_ERR_UNSUPPORTED, "only TYPE_FLOAT and TYPE_DOUBLE are supported");

    const bool dev = is_device_pointer(A->data) && is_device_pointer(B->data) && is_device_pointer(C->data);
    const void *dA = A->data, *dB = B->data;
    void* dC = C->data;
    void *bufA = nullptr, *bufB = nullptr, *bufC = nullptr;
    int64_t loA = 0, hiA = 0, loB = 0, hiB = 0, loC = 0, hiC = 0;
    auto cleanup = [&]() {
        if (bufA) (void)hipFreeAsync(bufA, stream);
        if (bufB) (void)hipFreeAsync(bufB, stream);
        if (bufC) (void)hipFreeAsync(bufC, stream);
    };
#define TRY_HIP(expr)                                                                         \
    do                                                                                        \
    {                                                                                         \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
        {                                                                                     \
            cleanup();                                                                        \
            return Status::error(MISTRA_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
        }                                                                                     \
    } while (0)

    if (!dev && !empty_tensor(C))
    {
        // host tensors: stage the touched span of each through HBM
        span(A, loA, hiA);
        span(B, loB, hiB);
        span(C, loC, hiC);
        const size_t nA = empty_tensor(A) ? 0 : size_t(hiA - loA + 1) * elem;
        const size_t nB = empty_tensor(B) ? 0 : size_t(hiB - loB + 1) * elem;
        const size_t nC = size_t(hiC - loC + 1) * elem;
        if (nA) TRY_HIP(hipMallocAsync(&bufA, nA, stream));
        if (nB) TRY_HIP(hipMallocAsync(&bufB, nB, stream));
        TRY_HIP(hipMallocAsync(&bufC, nC, stream));
        if (nA)
            TRY_HIP(hipMemcpyAsync(bufA, static_cast<const char*>(A->data) + loA * elem, nA, hipMemcpyHostToDevice,
                                   stream));
        if (nB)
            TRY_HIP(hipMemcpyAsync(bufB, static_cast<const char*>(B->data) + loB * elem, nB, hipMemcpyHostToDevice,
                                   stream));
        TRY_HIP(hipMemcpyAsync(bufC, static_cast<const char*>(C->data) + loC * elem, nC, hipMemcpyHostToDevice,
                               stream));
        dA = static_cast<const char*>(bufA) - loA * elem;
        dB = static_cast<const char*>(bufB) - loB * elem;
        dC = static_cast<char*>(bufC) - loC * elem;
    }

    if (!empty_tensor(C))
    {
        if (!plan)
        {
            // alpha == 0 or an empty operand: C = beta*C (src/stra_iface/3t/stra_mult.cxx:212-231)
            st = scale_c(C, dC, beta, stream);
        }
        else
        {
            Call call{dA, dB, dC, alpha, beta};
            if (plan->transposed) std::swap(call.A, call.B);
            st = run_plan(*plan, call, stream);
        }
        if (!st.ok())
        {
            cleanup();
            return st;
        }
    }

    if (bufC)
    {
        const size_t nC = size_t(hiC - loC + 1) * elem;
        TRY_HIP(hipMemcpyAsync(static_cast<char*>(C->data) + loC * elem, bufC, nC, hipMemcpyDeviceToHost, stream));
        cleanup();
        TRY_HIP(hipStreamSynchronize(stream));
    }
    else if (synchronous || (flags & MISTRA_FLAG_SYNC))
    {
        TRY_HIP(hipStreamSynchronize(stream));
    }
#undef TRY_HIP
    set_scalar_one(C);
    return Status();
}

}  // namespace mistra

// ---- C-ABI -----------------------------------------------------------------------------------------------

using mistra::fail;
using mistra::Status;

namespace tblis {
extern "C" {

void stra_tblis_tensor_mult(const tblis_comm*, const tblis_config*, const tblis_tensor* A, const label_type* idx_A,
                            const tblis_tensor* B, const label_type* idx_B, tblis_tensor* C, const label_type* idx_C)
{
    mistra::g_last_error.clear();
    Status s = mistra::tensor_mult(A, idx_A, B, idx_B, C, idx_C, 1, int(internal::impl), MISTRA_PREC_NATIVE, 0,
                                   nullptr, true);
    if (!s.ok()) fail(s);
}

void stra_tblis_tensor_mult_2level(const tblis_comm*, const tblis_config*, const tblis_tensor* A,
                                   const label_type* idx_A, const tblis_tensor* B, const label_type* idx_B,
                                   tblis_tensor* C, const label_type* idx_C)
{
    mistra::g_last_error.clear();
    Status s = mistra::tensor_mult(A, idx_A, B, idx_B, C, idx_C, 2, int(internal::impl), MISTRA_PREC_NATIVE, 0,
                                   nullptr, true);
    if (!s.ok()) fail(s);
}

void tblis_tensor_mult(const tblis_comm*, const tblis_config*, const tblis_tensor* A, const label_type* idx_A,
                       const tblis_tensor* B, const label_type* idx_B, tblis_tensor* C, const label_type* idx_C)
{
    // L=0 is classical for every impl (src/internal/3t/mult.cxx:705-823 never consults the Strassen impls)
    mistra::g_last_error.clear();
    Status s = mistra::tensor_mult(A, idx_A, B, idx_B, C, idx_C, 0, mistra::IMPL_ABC, MISTRA_PREC_NATIVE, 0, nullptr,
                                   true);
    if (!s.ok()) fail(s);
}

int mistra_tensor_mult(const tblis_tensor* A, const label_type* idx_A, const tblis_tensor* B, const label_type* idx_B,
                       tblis_tensor* C, const label_type* idx_C, int level, int impl, int precision, unsigned flags,
                       void* stream)
{
    mistra::g_last_error.clear();
    Status s = mistra::tensor_mult(A, idx_A, B, idx_B, C, idx_C, level, impl, precision, flags,
                                   static_cast<hipStream_t>(stream), false);
    return s.ok() ? 0 : fail(s);
}

const char* mistra_last_error(void) { return mistra::g_last_error.c_str(); }

int mistra_set_impl(int impl)
{
    if (impl < 0 || impl > 4) return fail(Status::error(MISTRA_ERR_ARG, "impl must be in 0..4"));
    internal::impl = static_cast<internal::impl_t>(impl);
    return 0;
}

int mistra_get_impl(void) { return int(internal::impl); }

int mistra_release_workspace(void)
{
    mistra::release_all();
    return 0;
}

int mistra_kernel_timing(int enable)
{
    std::lock_guard<std::mutex> lk(mistra::g_timing_mutex);
    mistra::g_timing = enable != 0;
    return 0;
}

int mistra_kernel_timing_read(double* total_ms, long long* launches, double* flops, int reset)
{
    std::lock_guard<std::mutex> lk(mistra::g_timing_mutex);
    double ms = 0.0, fl = 0.0;
    for (auto& t : mistra::g_pending)
    {
        hipError_t e = hipEventSynchronize(t.stop);
        if (e != hipSuccess) return fail(Status::error(MISTRA_ERR_HIP, hipGetErrorString(e)));
        float x = 0.f;
        e = hipEventElapsedTime(&x, t.start, t.stop);
        if (e != hipSuccess) return fail(Status::error(MISTRA_ERR_HIP, hipGetErrorString(e)));
        ms += double(x);
        fl += t.flops;
    }
    if (total_ms) *total_ms = ms;
    if (launches) *launches = (long long)mistra::g_pending.size();
    if (flops) *flops = fl;
    if (reset)
    {
        for (auto& t : mistra::g_pending)
        {
            (void)hipEventDestroy(t.start);
            (void)hipEventDestroy(t.stop);
        }
        mistra::g_pending.clear();
    }
    return 0;
}

int mistra_fill_uniform(int type, void* ptr, long long n, unsigned long long seed, int scale_exp_max, void* stream)
{
    if (type != TYPE_FLOAT && type != TYPE_DOUBLE) return fail(Status::error(MISTRA_ERR_ARG, "type"));
    hipError_t e = mistra::launch_fill_uniform(type == TYPE_DOUBLE ? 8 : 4, ptr, n, seed, scale_exp_max,
                                               static_cast<hipStream_t>(stream));
    return e == hipSuccess ? 0 : fail(Status::error(MISTRA_ERR_HIP, hipGetErrorString(e)));
}

int mistra_copy_bytes(void* dst, const void* src, long long nbytes, void* stream)
{
    hipError_t e = mistra::launch_copy_bytes(dst, src, nbytes, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? 0 : fail(Status::error(MISTRA_ERR_HIP, hipGetErrorString(e)));
}

int mistra_rel_error(int type, const void* X, const void* Y, long long n, double* out, void* stream)
{
    if (type != TYPE_FLOAT && type != TYPE_DOUBLE) return fail(Status::error(MISTRA_ERR_ARG, "type"));
    hipStream_t s = static_cast<hipStream_t>(stream);
    double* buf = nullptr;
    hipError_t e = hipMallocAsync(reinterpret_cast<void**>(&buf), sizeof(double) * (2 + 2 * 1024), s);
    if (e != hipSuccess) return fail(Status::error(MISTRA_ERR_NOMEM, hipGetErrorString(e)));
    e = mistra::launch_rel_error(type == TYPE_DOUBLE ? 8 : 4, X, Y, n, buf, s);
    double h[2] = {0, 0};
    if (e == hipSuccess) e = hipMemcpyAsync(h, buf, sizeof(h), hipMemcpyDeviceToHost, s);
    hipError_t f = hipFreeAsync(buf, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess) e = f;
    if (e != hipSuccess) return fail(Status::error(MISTRA_ERR_HIP, hipGetErrorString(e)));
    *out = h[1] > 0 ? std::sqrt(h[0] / h[1]) : std::sqrt(h[0]);
    return 0;
}

int mistra_batch_mult(int, const mistra_contraction*, unsigned, void*)
{
    return fail(Status::error(MISTRA_ERR_UNSUPPORTED, "batching not available in this build"));
}

int mistra_tensor_mult_backward(const tblis_tensor* A, const label_type* idx_A, const tblis_tensor* B,
                                const label_type* idx_B, const tblis_tensor* dC, const label_type* idx_C,
                                tblis_tensor* dA, tblis_tensor* dB, int level, int impl, int precision, void* stream)
{
    // The adjoints of a contraction are contractions over the same labels with the roles exchanged:
    // dA[idx_A] = alpha * dC[idx_C] . B[idx_B] (summed over the labels of B not in A), and
    // dB[idx_B] = alpha * A[idx_A] . dC[idx_C]; alpha = scalar(A) * scalar(B) * scalar(dC).
    mistra::g_last_error.clear();
    if (!A || !B || !dC) return fail(Status::error(MISTRA_ERR_ARG, "null tensor"));
    hipStream_t s = static_cast<hipStream_t>(stream);
    const double aA = mistra::scalar_as_double(A), aB = mistra::scalar_as_double(B),
                 aC = mistra::scalar_as_double(dC);
    if (dA)
    {
        tblis_tensor X = *dC, Y = *B;
        mistra::set_scalar(&X, aA * aC);
        Status st = mistra::tensor_mult(&X, idx_C, &Y, idx_B, dA, idx_A, level, impl, precision, 0, s, false);
        if (!st.ok()) return fail(st);
    }
    if (dB)
    {
        tblis_tensor X = *A, Y = *dC;
        mistra::set_scalar(&Y, aB * aC);
        Status st = mistra::tensor_mult(&X, idx_A, &Y, idx_C, dB, idx_B, level, impl, precision, 0, s, false);
        if (!st.ok()) return fail(st);
    }
    return 0;
}

int mistra_autotune(const tblis_tensor*, const label_type*, const tblis_tensor*, const label_type*,
                    const tblis_tensor*, const label_type*, int, int*, int*, int*, void*)
{
    return fail(Status::error(MISTRA_ERR_UNSUPPORTED, "autotuner not available in this build"));
}

}  // extern "C"
}  // namespace tblis
