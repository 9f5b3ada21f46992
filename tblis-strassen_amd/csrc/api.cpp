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
#include <cstdlib>
#include <cstring>
#include <list>
#include <map>
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

Status merged_dims_c(const tblis::tblis_tensor* t, const label_type* idx, StridedDims& d);

// C = beta*C over the elements that C[idx_C] addresses: a repeated label addresses a diagonal
// (src/stra_iface/3t/stra_mult.cxx:47-50 applies diagonal() to C before :212-231 scales it), so the
// elements off that diagonal keep their values.
Status scale_c(const tblis::tblis_tensor* C, const label_type* idx_C, void* data, double beta, hipStream_t s)
{
    StridedDims d;
    Status st = merged_dims_c(C, idx_C, d);
    if (!st.ok()) return st;
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

// Digit slices of the Ozaki mode: 5 (15 products, about 40 bits) unless MISTRA_OZAKI_SLICES says 2..7.
int ozaki_slices()
{
    if (const int o = g_dbg_ozaki_slices.load()) return o;
    static const int n = [] {
        const char* v = std::getenv("MISTRA_OZAKI_SLICES");
        const int x = v ? std::atoi(v) : 5;
        return x >= 2 && x <= 7 ? x : 5;
    }();
    return n;
}

// Digit slices of the INT8 Ozaki mode: 6 (21 products, about 41 bits) unless MISTRA_OZAKI_I8_SLICES
// says 2..6.
int ozaki_i8_slices()
{
    if (const int o = g_dbg_i8_slices.load()) return o;
    static const int n = [] {
        const char* v = std::getenv("MISTRA_OZAKI_I8_SLICES");
        const int x = v ? std::atoi(v) : 6;
        return x >= 2 && x <= 6 ? x : 6;
    }();
    return n;
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

namespace {

struct Dim
{
    char label;
    int64_t len, stride;
};

// Repeated labels of one tensor address its diagonal: one dimension with the summed stride
// (src/util/tensor.hpp:305-342).
Status merged_dims(const tblis::tblis_tensor* t, const label_type* idx, std::vector<Dim>& out)
{
    out.clear();
    for (unsigned i = 0; i < t->ndim; i++)
    {
        bool found = false;
        for (Dim& d : out)
            if (d.label == idx[i])
            {
                if (d.len != t->len[i]) return Status::error(MISTRA_ERR_ARG, "repeated label with different lengths");
                d.stride += t->stride[i];
                found = true;
            }
        if (!found) out.push_back(Dim{idx[i], t->len[i], t->stride[i]});
    }
    if (out.size() > 8) return Status::error(MISTRA_ERR_UNSUPPORTED, "more than 8 distinct labels in one tensor");
    return Status();
}

Status merged_dims_c(const tblis::tblis_tensor* t, const label_type* idx, StridedDims& d)
{
    std::vector<Dim> v;
    Status st = merged_dims(t, idx, v);
    if (!st.ok()) return st;
    std::memset(&d, 0, sizeof(d));
    d.ndim = int(v.size());
    for (size_t i = 0; i < v.size(); i++)
    {
        d.len[i] = v[i].len;
        d.stride[i] = v[i].stride;
    }
    return Status();
}

bool has_label(const std::vector<Dim>& v, char l)
{
    for (const Dim& d : v)
        if (d.label == l) return true;
    return false;
}

// A dense (first index fastest) temporary with the given labels and lengths, as a tblis_tensor.
struct Temp
{
    std::string labels;
    std::vector<ptrdiff_t> len, stride;
    void* data = nullptr;
    tblis::tblis_tensor t;
};

void make_temp(Temp& x, const std::vector<Dim>& dims, int type)
{
    x.labels.clear();
    x.len.clear();
    x.stride.clear();
    int64_t st = 1;
    for (const Dim& d : dims)
    {
        x.labels.push_back(d.label);
        x.len.push_back(d.len);
        x.stride.push_back(st);
        st *= d.len;
    }
    std::memset(&x.t, 0, sizeof(x.t));
    x.t.type = tblis::type_t(type);
    x.t.ndim = unsigned(dims.size());
    x.t.len = x.len.data();
    x.t.stride = x.stride.data();
}

int64_t volume(const std::vector<Dim>& dims)
{
    int64_t n = 1;
    for (const Dim& d : dims) n *= d.len;
    return n;
}

}  // namespace

// General index classes, the classical fallback of the reference (stra_mult_blas,
// src/stra_internal/3t/stra_mult.cxx:1311-1382): labels only in A (only in B) are summed out of A
// (B) into dense temporaries, the remaining pure contraction runs through the Strassen path with the
// requested level and impl, and labels only in C receive the result by broadcast: C = R + beta*C.
Status general_mult(const tblis::tblis_tensor* A, const label_type* idx_A, const void* dA,
                    const tblis::tblis_tensor* B, const label_type* idx_B, const void* dB,
                    const tblis::tblis_tensor* C, const label_type* idx_C, void* dC, int level, int impl,
                    int precision, unsigned flags, double alpha, double beta, int elem, hipStream_t stream)
{
    std::vector<Dim> a, b, c;
    Status st = merged_dims(A, idx_A, a);
    if (st.ok()) st = merged_dims(B, idx_B, b);
    if (st.ok()) st = merged_dims(C, idx_C, c);
    if (!st.ok()) return st;
    std::vector<void*> temps;
    auto release = [&]() {
        for (void* p : temps) (void)hipFreeAsync(p, stream);
    };
    auto alloc = [&](size_t bytes, void** out) -> Status {
        hipError_t e = mistra::stream_alloc(out, bytes ? bytes : 1, stream);
        if (e != hipSuccess) return Status::error(MISTRA_ERR_NOMEM, std::string("general path: ") + hipGetErrorString(e));
        temps.push_back(*out);
        return Status();
    };

    // operand X with its one-operand labels summed out
    auto prepare = [&](const std::vector<Dim>& x, const std::vector<Dim>& o1, const std::vector<Dim>& o2,
                       const void* data, Temp& out) -> Status {
        std::vector<Dim> keep, red;
        for (const Dim& d : x) (has_label(o1, d.label) || has_label(o2, d.label) ? keep : red).push_back(d);
        make_temp(out, keep, C->type);
        if (red.empty())
        {
            // no reduction: view the operand itself with its merged labels
            for (size_t i = 0; i < keep.size(); i++) out.stride[i] = keep[i].stride;
            out.data = const_cast<void*>(data);
        }
        else
        {
            Status s2 = alloc(size_t(volume(keep)) * elem, &out.data);
            if (!s2.ok()) return s2;
            ReduceArgs ra;
            std::memset(&ra, 0, sizeof(ra));
            ra.out.ndim = int(keep.size());
            for (size_t i = 0; i < keep.size(); i++)
            {
                ra.out.len[i] = keep[i].len;
                ra.out.stride[i] = keep[i].stride;
            }
            ra.red.ndim = int(red.size());
            for (size_t i = 0; i < red.size(); i++)
            {
                ra.red.len[i] = red[i].len;
                ra.red.stride[i] = red[i].stride;
            }
            hipError_t e = launch_reduce(elem, data, out.data, ra, stream);
            if (e != hipSuccess) return Status::error(MISTRA_ERR_HIP, hipGetErrorString(e));
        }
        out.t.data = out.data;
        return Status();
    };
    Temp ta, tb, tc;
    st = prepare(a, b, c, dA, ta);
    if (st.ok()) st = prepare(b, a, c, dB, tb);
    if (!st.ok())
    {
        release();
        return st;
    }
    set_scalar(&ta.t, alpha);
    set_scalar(&tb.t, 1.0);

    std::vector<Dim> keep;
    bool conly = false;
    for (const Dim& d : c)
    {
        if (has_label(a, d.label) || has_label(b, d.label))
            keep.push_back(d);
        else
            conly = true;
    }
    make_temp(tc, keep, C->type);
    if (!conly)
    {
        for (size_t i = 0; i < keep.size(); i++) tc.stride[i] = keep[i].stride;
        tc.t.data = dC;
        set_scalar(&tc.t, beta);
        st = tensor_mult(&ta.t, ta.labels.c_str(), &tb.t, tb.labels.c_str(), &tc.t, tc.labels.c_str(), level, impl,
                         precision, flags & ~unsigned(MISTRA_FLAG_SYNC), stream, false);
    }
    else
    {
        st = alloc(size_t(volume(keep)) * elem, &tc.data);
        if (st.ok())
        {
            tc.t.data = tc.data;
            set_scalar(&tc.t, 0.0);
            st = tensor_mult(&ta.t, ta.labels.c_str(), &tb.t, tb.labels.c_str(), &tc.t, tc.labels.c_str(), level,
                             impl, precision, flags & ~unsigned(MISTRA_FLAG_SYNC), stream, false);
        }
        if (st.ok())
        {
            BcastArgs ba;
            std::memset(&ba, 0, sizeof(ba));
            ba.c.ndim = int(c.size());
            for (size_t i = 0; i < c.size(); i++)
            {
                ba.c.len[i] = c[i].len;
                ba.c.stride[i] = c[i].stride;
                for (size_t j = 0; j < keep.size(); j++)
                    if (keep[j].label == c[i].label) ba.r_stride[i] = tc.stride[j];
            }
            ba.beta = beta;
            hipError_t e = launch_bcast_add(elem, tc.data, dC, ba, stream);
            if (e != hipSuccess) st = Status::error(MISTRA_ERR_HIP, hipGetErrorString(e));
        }
    }
    release();
    return st;
}

namespace {

// FP32 tensors in a low-precision mode: the modes form their operand sums, split / scale computation and
// combine in FP64, so A, B and C are copied to FP64 (the touched span of each, same strides), the FP64
// contraction runs on the copies, and the addressed elements of C are rounded back to FP32.  Device
// tensors only.
Status lowp_float(const tblis::tblis_tensor* A, const label_type* idx_A, const tblis::tblis_tensor* B,
                  const label_type* idx_B, tblis::tblis_tensor* C, const label_type* idx_C, int level, int impl,
                  int precision, unsigned flags, hipStream_t stream, bool synchronous, int tile)
{
    if (A->type != tblis::TYPE_FLOAT || B->type != tblis::TYPE_FLOAT)
        return Status::error(MISTRA_ERR_ARG, "A, B and C must have the same type");
    if (!is_device_pointer(A->data) || !is_device_pointer(B->data) || !is_device_pointer(C->data))
        return Status::error(MISTRA_ERR_UNSUPPORTED, "low-precision modes on FP32 tensors need device memory");
    StridedDims cd;
    Status st = merged_dims_c(C, idx_C, cd);
    if (!st.ok()) return st;
    const tblis::tblis_tensor* src[3] = {A, B, C};
    tblis::tblis_tensor t64[3];
    void* buf[3] = {nullptr, nullptr, nullptr};
    auto release = [&]() {
        for (void* b : buf)
            if (b) (void)hipFreeAsync(b, stream);
    };
    for (int i = 0; i < 3; i++)
    {
        t64[i] = *src[i];
        t64[i].type = tblis::TYPE_DOUBLE;
        set_scalar(&t64[i], scalar_as_double(src[i]));
        if (empty_tensor(src[i])) continue;
        int64_t lo = 0, hi = 0;
        span(src[i], lo, hi);
        const int64_t n = hi - lo + 1;
        hipError_t e = mistra::stream_alloc(&buf[i], size_t(n) * 8, stream);
        if (e == hipSuccess)
            e = launch_f32_to_f64(static_cast<const float*>(src[i]->data) + lo, static_cast<double*>(buf[i]), n,
                                  stream);
        if (e != hipSuccess)
        {
            release();
            return Status::error(e == hipErrorOutOfMemory ? MISTRA_ERR_NOMEM : MISTRA_ERR_HIP,
                                 std::string("FP64 copies of FP32 tensors: ") + hipGetErrorString(e));
        }
        t64[i].data = static_cast<double*>(buf[i]) - lo;
    }
    st = tensor_mult(&t64[0], idx_A, &t64[1], idx_B, &t64[2], idx_C, level, impl, precision,
                     flags & ~unsigned(MISTRA_FLAG_SYNC), stream, false, tile);
    if (st.ok() && buf[2])
    {
        hipError_t e = launch_f64_to_f32_strided(static_cast<const double*>(t64[2].data), static_cast<float*>(C->data),
                                                 cd, stream);
        if (e != hipSuccess) st = Status::error(MISTRA_ERR_HIP, hipGetErrorString(e));
    }
    release();
    if (st.ok() && (synchronous || (flags & MISTRA_FLAG_SYNC)))
    {
        hipError_t e = hipStreamSynchronize(stream);
        if (e != hipSuccess) st = Status::error(MISTRA_ERR_HIP, hipGetErrorString(e));
    }
    if (st.ok()) set_scalar_one(C);
    return st;
}

}  // namespace

Status tensor_mult(const tblis::tblis_tensor* A, const label_type* idx_A, const tblis::tblis_tensor* B,
                   const label_type* idx_B, tblis::tblis_tensor* C, const label_type* idx_C, int level, int impl,
                   int precision, unsigned flags, hipStream_t stream, bool synchronous, int tile)
{
    Status st = validate(A, idx_A, B, idx_B, C, idx_C);
    if (!st.ok()) return st;
    if (precision < MISTRA_PREC_NATIVE || precision > MISTRA_PREC_OZAKI_INT8)
        return Status::error(MISTRA_ERR_ARG, "unknown precision mode");
    if (precision != MISTRA_PREC_NATIVE && C->type == tblis::TYPE_FLOAT)
        return lowp_float(A, idx_A, B, idx_B, C, idx_C, level, impl, precision, flags, stream, synchronous, tile);
    if (precision != MISTRA_PREC_NATIVE && C->type != tblis::TYPE_DOUBLE)
        return Status::error(MISTRA_ERR_UNSUPPORTED, "low-precision modes take FP32 or FP64 tensors");
    if (level < 0 && impl != MISTRA_IMPL_AUTO)
        level = best_level(A, idx_A, B, idx_B, C, idx_C, impl);
    else if (level < 0 || impl == MISTRA_IMPL_AUTO)
    {
        TuneChoice ch;
        st = autotune(A, idx_A, B, idx_B, C, idx_C, false, stream, ch);
        if (!st.ok()) return st;
        if (level < 0) level = ch.level;
        if (impl == MISTRA_IMPL_AUTO) impl = ch.impl;
        if (tile < 0) tile = ch.tile;
    }
    if (tile < 0 && g_dbg_tile.load() >= 0) tile = g_dbg_tile.load();
    if (tile >= kNumTiles) return Status::error(MISTRA_ERR_ARG, "unknown register tile");
    if (level > 2) return Status::error(MISTRA_ERR_ARG, "level must be 0, 1 or 2");
    if (impl < 0 || impl > 4) return Status::error(MISTRA_ERR_ARG, "impl must be in 0..4");
    // low-precision leaves always stage the operand sums (the Naive tables of the plan)
    if (precision != MISTRA_PREC_NATIVE) impl = IMPL_NAIVE;

    const double alpha = scalar_as_double(A) * scalar_as_double(B);
    const double beta = scalar_as_double(C);

    // A caller may record the call into a HIP graph: then the plan must exist already (building one
    // uploads its tables synchronously) and stays resident for the graph's replays; host tensors (staged
    // with copies from pageable memory) cannot be captured.
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(stream, &cap) != hipSuccess)
    {
        (void)hipGetLastError();
        cap = hipStreamCaptureStatusNone;
    }
    const bool capturing = cap != hipStreamCaptureStatusNone;

    std::shared_ptr<Plan> plan;
    if (!empty_tensor(C) && alpha != 0.0 && !empty_tensor(A) && !empty_tensor(B))
    {
        st = get_plan(A, idx_A, B, idx_B, C, idx_C, level, impl, tile, flags, plan, capturing);
        if (!st.ok()) return st;
    }
    else if (A->type != C->type || B->type != C->type)
        return Status::error(MISTRA_ERR_ARG, "A, B and C must have the same type");

    const int elem = C->type == tblis::TYPE_DOUBLE ? 8 : (C->type == tblis::TYPE_FLOAT ? 4 : 0);
    if (!elem) return Status::error(MISTRA_ERR_UNSUPPORTED, "only TYPE_FLOAT and TYPE_DOUBLE are supported");

    const bool dev = is_device_pointer(A->data) && is_device_pointer(B->data) && is_device_pointer(C->data);
    if (capturing && !dev)
        return Status::error(MISTRA_ERR_UNSUPPORTED, "host tensors cannot be used inside a HIP graph capture");
    if (capturing && synchronous)
        return Status::error(MISTRA_ERR_UNSUPPORTED, "a synchronous call cannot be recorded into a HIP graph");
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
        if (nA) TRY_HIP(mistra::stream_alloc(&bufA, nA, stream));
        if (nB) TRY_HIP(mistra::stream_alloc(&bufB, nB, stream));
        TRY_HIP(mistra::stream_alloc(&bufC, nC, stream));
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
            st = scale_c(C, idx_C, dC, beta, stream);
        }
        else
        {
            if (plan->general)
                st = general_mult(A, idx_A, dA, B, idx_B, dB, C, idx_C, dC, level, impl, precision, flags, alpha, beta,
                                  elem, stream);
            else
            {
                Call call{dA, dB, dC, alpha, beta};
                call.flags = flags;
                if (plan->transposed) std::swap(call.A, call.B);
                if (precision == MISTRA_PREC_NATIVE)
                    st = run_plan(*plan, call, stream);
                else
                    st = run_plan_lowp(*plan, call, precision,
                                       precision == MISTRA_PREC_BF16X3       ? 2
                                       : precision == MISTRA_PREC_OZAKI_INT8 ? ozaki_i8_slices()
                                                                             : ozaki_slices(),
                                       stream);
            }
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
        const Status ws = stage_wait_status();
        if (!ws.ok()) return ws;
    }
#undef TRY_HIP
    set_scalar_one(C);
    return Status();
}

namespace {

// Plans of the allocation-free entry point: built on the host, their tables bound to the device's
// scratch block (the image is copied there when a call needs a plan other than the one resident).
struct ScratchPlan
{
    std::shared_ptr<Plan> plan;  // device tables point into the scratch block (dmem stays null)
    PlanImage img;
    size_t table_bytes = 0;      // the image, rounded up to 256 bytes: workspace starts there
};
std::mutex g_noalloc_mutex;
std::list<std::pair<std::string, std::shared_ptr<ScratchPlan>>> g_noalloc_plans;  // LRU, most recent first
constexpr size_t kMaxScratchPlans = 64;
std::map<int, std::shared_ptr<ScratchPlan>> g_resident;  // device -> the plan whose tables are in its scratch

}  // namespace

// mistra_tensor_mult without device allocation: the plan's tables and the staged workspaces of its
// launches live in the fixed scratch block (noalloc.cpp), nothing is allocated or freed on the device.
// Native precision, explicit level and impl, device tensors, pure contractions (no A-only / B-only /
// C-only labels), not under graph capture.  Staged ABC falls back to the fused form when its sums do
// not fit; AB and Naive need their staged products to fit and report an error otherwise.
Status tensor_mult_noalloc(const tblis::tblis_tensor* A, const label_type* idx_A, const tblis::tblis_tensor* B,
                           const label_type* idx_B, tblis::tblis_tensor* C, const label_type* idx_C, int level,
                           int impl, int precision, unsigned flags, hipStream_t stream)
{
    Status st = validate(A, idx_A, B, idx_B, C, idx_C);
    if (!st.ok()) return st;
    if (precision != MISTRA_PREC_NATIVE)
        return Status::error(MISTRA_ERR_UNSUPPORTED, "the allocation-free entry point runs native precision only");
    if (level < 0 || level > 2 || impl < 0 || impl > 4)
        return Status::error(MISTRA_ERR_ARG, "the allocation-free entry point needs level 0..2 and impl 0..4");
    if (A->type != C->type || B->type != C->type)
        return Status::error(MISTRA_ERR_ARG, "A, B and C must have the same type");
    const int elem = C->type == tblis::TYPE_DOUBLE ? 8 : (C->type == tblis::TYPE_FLOAT ? 4 : 0);
    if (!elem) return Status::error(MISTRA_ERR_UNSUPPORTED, "only TYPE_FLOAT and TYPE_DOUBLE are supported");
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(stream, &cap) != hipSuccess)
    {
        (void)hipGetLastError();
        cap = hipStreamCaptureStatusNone;
    }
    if (cap != hipStreamCaptureStatusNone)
        return Status::error(MISTRA_ERR_UNSUPPORTED,
                             "the allocation-free entry point cannot be captured (its scratch contents change "
                             "from call to call)");
    if (!is_device_pointer(A->data) || !is_device_pointer(B->data) || !is_device_pointer(C->data))
        return Status::error(MISTRA_ERR_UNSUPPORTED, "the allocation-free entry point takes device tensors only");
    const double alpha = scalar_as_double(A) * scalar_as_double(B);
    const double beta = scalar_as_double(C);
    if (empty_tensor(C))
    {
        set_scalar_one(C);
        return Status();
    }
    if (alpha == 0.0 || empty_tensor(A) || empty_tensor(B))
    {
        st = scale_c(C, idx_C, C->data, beta, stream);
        if (st.ok() && (flags & MISTRA_FLAG_SYNC))
        {
            const hipError_t e = hipStreamSynchronize(stream);
            if (e != hipSuccess) st = Status::error(MISTRA_ERR_HIP, hipGetErrorString(e));
        }
        if (st.ok()) set_scalar_one(C);
        return st;
    }

    void* base = nullptr;
    size_t cap_bytes = 0;
    st = noalloc_scratch(&base, &cap_bytes);
    if (!st.ok()) return st;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return Status::error(MISTRA_ERR_HIP, hipGetErrorString(e));
    const std::string key = plan_key(A, idx_A, B, idx_B, C, idx_C, level, impl, -1, flags, dev);

    std::lock_guard<std::mutex> lk(g_noalloc_mutex);
    std::shared_ptr<ScratchPlan> sp;
    for (auto it = g_noalloc_plans.begin(); it != g_noalloc_plans.end(); ++it)
        if (it->first == key)
        {
            sp = it->second;
            g_noalloc_plans.splice(g_noalloc_plans.begin(), g_noalloc_plans, it);
            break;
        }
    if (!sp)
    {
        auto n = std::make_shared<ScratchPlan>();
        n->plan = std::make_shared<Plan>();
        st = build_plan_host(A, idx_A, B, idx_B, C, idx_C, level, impl, -1, flags, *n->plan);
        if (!st.ok()) return st;
        if (n->plan->general)
            return Status::error(MISTRA_ERR_UNSUPPORTED,
                                 "the allocation-free entry point takes pure contractions only (no labels that "
                                 "appear in A, B or C alone)");
        build_image(*n->plan, n->img);
        n->table_bytes = (n->img.bytes.size() + 255) & ~size_t(255);
        if (n->table_bytes > cap_bytes)
            return Status::error(MISTRA_ERR_NOMEM, "the plan's tables need " + std::to_string(n->img.bytes.size()) +
                                                       " bytes, more than the allocation-free scratch block of " +
                                                       std::to_string(cap_bytes) + " bytes");
        bind_image(*n->plan, n->img, base);
        g_noalloc_plans.emplace_front(key, n);
        while (g_noalloc_plans.size() > kMaxScratchPlans) g_noalloc_plans.pop_back();
        sp = n;
    }
    const Plan& p = *sp->plan;
    // staged products of AB / Naive must fit beside the tables (ABC's staged sums fall back to the fused form)
    const size_t room = cap_bytes - sp->table_bytes;
    if (p.core && (p.impl == IMPL_AB || p.impl == IMPL_NAIVE))
    {
        const size_t q = size_t(p.nprod) * size_t(p.elem), up = 255;
        const size_t need = p.impl == IMPL_AB
                                ? ((size_t(p.nbatch) * q * size_t(p.ms) * size_t(p.ns) + up) & ~up)
                                : ((q * size_t(p.ms) * size_t(p.ks) + up) & ~up) +
                                      ((q * size_t(p.ks) * size_t(p.ns) + up) & ~up) +
                                      ((q * size_t(p.ms) * size_t(p.ns) + up) & ~up);
        if (need > room)
            return Status::error(MISTRA_ERR_NOMEM,
                                 std::string(p.impl == IMPL_AB ? "AB" : "Naive") + " needs " + std::to_string(need) +
                                     " bytes of staged products, more than the " + std::to_string(room) +
                                     " bytes the allocation-free scratch block has left beside the plan's tables");
    }
    // the tables of this plan into the scratch block, unless they are there already
    auto res = g_resident.find(dev);
    if (res == g_resident.end() || res->second != sp)
    {
        e = hipMemcpyAsync(base, sp->img.bytes.data(), sp->img.bytes.size(), hipMemcpyHostToDevice, stream);
        if (e != hipSuccess) return Status::error(MISTRA_ERR_HIP, std::string("scratch upload: ") + hipGetErrorString(e));
        g_resident[dev] = sp;  // keeps the image alive while it is the resident one
    }
    Arena ws;
    ws.base = static_cast<char*>(base) + sp->table_bytes;
    ws.size = room;
    Call call{A->data, B->data, C->data, alpha, beta, &ws};
    call.flags = flags;
    if (p.transposed) std::swap(call.A, call.B);
    st = run_plan(p, call, stream);
    if (st.ok() && (flags & MISTRA_FLAG_SYNC))
    {
        e = hipStreamSynchronize(stream);
        st = e != hipSuccess ? Status::error(MISTRA_ERR_HIP, hipGetErrorString(e)) : stage_wait_status();
    }
    if (st.ok()) set_scalar_one(C);
    return st;
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

}  // extern "C"
}  // namespace tblis

namespace mistra {
namespace {

// The matrix entry points are contractions C[ij] = A[ik] B[kj] with two-dimensional views
// (src/stra_iface/3m/stra_mult.cxx: 1-level Strassen, impl by symbol).
void matrix_mult(const tblis::tblis_matrix* A, const tblis::tblis_matrix* B, tblis::tblis_matrix* C, int impl)
{
    g_last_error.clear();
    if (!A || !B || !C)
    {
        fail(Status::error(MISTRA_ERR_ARG, "null matrix"));
        return;
    }
    if (A->m != C->m || B->n != C->n || A->n != B->m)
    {
        fail(Status::error(MISTRA_ERR_ARG, "matrix dimensions do not conform"));
        return;
    }
    if (A->conj || B->conj || C->conj)
    {
        fail(Status::error(MISTRA_ERR_UNSUPPORTED, "conjugation is not supported"));
        return;
    }
    ptrdiff_t la[2] = {A->m, A->n}, sa[2] = {A->rs, A->cs};
    ptrdiff_t lb[2] = {B->m, B->n}, sb[2] = {B->rs, B->cs};
    ptrdiff_t lc[2] = {C->m, C->n}, sc[2] = {C->rs, C->cs};
    tblis::tblis_tensor tA, tB, tC;
    for (auto* x : {&tA, &tB, &tC}) std::memset(x, 0, sizeof(*x));
    auto fill = [](tblis::tblis_tensor& t, const tblis::tblis_matrix* m, ptrdiff_t* len, ptrdiff_t* str) {
        t.type = m->type;
        std::memcpy(t.scalar, m->scalar, sizeof(t.scalar));
        t.data = m->data;
        t.ndim = 2;
        t.len = len;
        t.stride = str;
    };
    fill(tA, A, la, sa);
    fill(tB, B, lb, sb);
    fill(tC, C, lc, sc);
    Status st = tensor_mult(&tA, "ik", &tB, "kj", &tC, "ij", 1, impl, MISTRA_PREC_NATIVE, 0, nullptr, true);
    if (!st.ok())
    {
        fail(st);
        return;
    }
    std::memcpy(C->scalar, tC.scalar, sizeof(C->scalar));
    C->conj = 0;
}

}  // namespace
}  // namespace mistra

namespace tblis {
extern "C" {

void stra_tblis_matrix_mult(const tblis_comm*, const tblis_config*, const tblis_matrix* A, const tblis_matrix* B,
                            tblis_matrix* C)
{
    mistra::matrix_mult(A, B, C, mistra::IMPL_ABC);
}

void stra_tblis_matrix_mult_naive(const tblis_comm*, const tblis_config*, const tblis_matrix* A,
                                  const tblis_matrix* B, tblis_matrix* C)
{
    mistra::matrix_mult(A, B, C, mistra::IMPL_NAIVE);
}

void stra_tblis_matrix_mult_ab(const tblis_comm*, const tblis_config*, const tblis_matrix* A, const tblis_matrix* B,
                               tblis_matrix* C)
{
    mistra::matrix_mult(A, B, C, mistra::IMPL_AB);
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

int mistra_tensor_mult_noalloc(const tblis_tensor* A, const label_type* idx_A, const tblis_tensor* B,
                               const label_type* idx_B, tblis_tensor* C, const label_type* idx_C, int level, int impl,
                               int precision, unsigned flags, void* stream)
{
    mistra::g_last_error.clear();
    Status s = mistra::tensor_mult_noalloc(A, idx_A, B, idx_B, C, idx_C, level, impl, precision, flags,
                                           static_cast<hipStream_t>(stream));
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
    mistra::release_batch_graphs();
    mistra::release_all();
    mistra::trim_stream_pool();
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
    hipError_t e = mistra::stream_alloc(reinterpret_cast<void**>(&buf), sizeof(double) * (2 + 2 * 1024), s);
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

int mistra_batch_mult(int count, const mistra_contraction* entries, unsigned flags, void* stream)
{
    mistra::g_last_error.clear();
    Status st = mistra::batch_mult(count, entries, flags, static_cast<hipStream_t>(stream));
    return st.ok() ? 0 : fail(st);
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

int mistra_autotune(const tblis_tensor* A, const label_type* idx_A, const tblis_tensor* B, const label_type* idx_B,
                    const tblis_tensor* C, const label_type* idx_C, int measure, int* level_out, int* impl_out,
                    int* tile_out, void* stream)
{
    mistra::g_last_error.clear();
    mistra::TuneChoice ch;
    Status st = mistra::autotune(A, idx_A, B, idx_B, C, idx_C, measure != 0, static_cast<hipStream_t>(stream), ch);
    if (!st.ok()) return fail(st);
    if (level_out) *level_out = ch.level;
    if (impl_out) *impl_out = ch.impl;
    if (tile_out) *tile_out = ch.tile;
    return 0;
}

}  // extern "C"
}  // namespace tblis
