// Autotuner: level, impl and register tile per contraction shape -- the GPU counterpart of the
// reference's static per-ISA blocksize table and runtime kernel selection (src/configs/configs.hpp:89-268,
// src/internal/3t/mult.cxx:26 for the impl global).  Two modes:
//   * analytic: a time model of the launches a plan makes, from leaf-kernel rates measured on MI355X;
//   * measured: every candidate timed on the device against a scratch copy of C, best kept per shape.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <sstream>

#include "driver.hpp"

namespace mistra {

namespace {

std::mutex g_tune_mutex;
std::map<std::string, TuneChoice> g_measured;
std::map<std::string, int> g_best_level;  // modelled level per (shape, impl)
constexpr size_t kMaxTuneEntries = 1 << 16;

std::string shape_key(const tblis::tblis_tensor* A, const label_type* idx_A, const tblis::tblis_tensor* B,
                      const label_type* idx_B, const tblis::tblis_tensor* C, const label_type* idx_C)
{
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::ostringstream k;
    k << dev << '|' << int(C->type);
    for (auto t : {std::make_pair(A, idx_A), std::make_pair(B, idx_B), std::make_pair(static_cast<const tblis::tblis_tensor*>(C), idx_C)})
    {
        k << '|';
        for (unsigned i = 0; i < t.first->ndim; i++)
            k << t.second[i] << ':' << t.first->len[i] << ':' << t.first->stride[i] << ',';
    }
    return k.str();
}

// Leaf-kernel rates (TFLOP/s) measured on MI355X with the 8192^3 FP64 workload, by register tile and
// number of sum terms (1, 2, 4).  Single-term launches over K-fast operands (K-tiles straight into LDS)
// run at 72.5 - 7645 / ks with 128 x 128 tiles: 71.6 at a leaf depth of 8192 (the classical kernel),
// 68.8 at 2048 (the L=2 Naive dense launch) -- the per-tile start and end weigh more on shorter K loops.
// FP32 runs the single-term kernels at about 1.8 times the rate (FP32 MFMA has twice the FP64 rate); the
// multi-term loaders' vector work is the same in both types and does not overlap the MFMAs, so FP32
// multi-term launches gain less: 2 terms measured 87 TFLOP/s (1.5x) at m = n = k = 4096 (c5 L1), 4 terms
// taken at 1.3x.
// The rectangular tiles (single-term launches only; rates net of the wave quantisation launch_s adds)
// against 128 x 128: FP64 about 1.5% above it (64 x 128) or level (128 x 64) when the epilogue stores the
// product (Naive's dense GEMM at 8192^3, as the benchmark times it; from one process to the next the three
// tiles of that launch trade places by up to 4%, each of them the fastest in some run), with the multi-C
// epilogue (the classical launch) 2.2% below it
// (128 x 64) and, for 64 x 128, 0.98 + 45 / ks of it: 1.4% below at a leaf depth of 8192, 2.4% above at
// 1024 (m = n = 16384, k = 1024), where its three resident workgroups per CU hide more of each tile's
// start and end; 2% above it in the register-staged form (an operand that is not K-fast); FP32 (classical
// at m = n = k = 4096) 1.23x / 1.16x with both operands K-fast, 1.04x / 1.03x otherwise.
double leaf_rate(int elem, int tile, int terms, bool dma, int64_t ks, bool store)
{
    static const double r128[3] = {61.0, 57.5, 46.0}, r64[3] = {57.0, 46.0, 38.0}, f32[3] = {1.8, 1.5, 1.3};
    const int ti = terms <= 1 ? 0 : (terms == 2 ? 1 : 2);
    const double dma128 = 72.5 - 7645.0 / double(std::max<int64_t>(ks, 512));
    double r = tile == TILE_64 ? r64[ti] : r128[ti];
    if (tile != TILE_64 && ti == 0 && dma) r = std::max(r, dma128);
    if (tile_is_rect(tile) && ti == 0)
    {
        const bool wide = tile == TILE_64x128;
        if (elem == 4) return f32[0] * r * (dma ? (wide ? 1.23 : 1.16) : (wide ? 1.037 : 1.034));
        const double wide_multi_c = 0.98 + 45.0 / double(std::max<int64_t>(ks, 1024));
        return r * (!dma ? 1.02 : store ? (wide ? 1.015 : 1.003) : (wide ? wide_multi_c : 0.978));
    }
    return elem == 8 ? r : f32[ti] * r;
}

// Share of an ABC stage launch that the multi-C epilogue costs (C targets per product over the leaf
// depth): the stages of an ABC launch sequence end all their workgroups' read-modify-writes of C at once.
// Measured at m = n = k = 8192: -9% at L=2 (about 3 targets per product, 2048-deep leaves), nothing
// visible at L=1 (1.7 targets, 4096 deep); with the rectangular tiles -21% and -3% (fewer columns share
// each row's C read).
double epilogue_penalty(int tile, double targets_per_product, int64_t ks)
{
    const double w = tile_is_rect(tile) ? 160.0 : 80.0;
    return std::max(0.0, std::min(0.3, w * targets_per_product / double(std::max<int64_t>(ks, 1)) - 0.033));
}

// Resident workgroups per CU (LDS and registers of the kernel variants): 128 x 128 two, the rectangular
// tiles three, 64 x 64 four.
int64_t resident_slots(int tile, int cus = 256)
{
    return int64_t(tile == TILE_128 ? 2 : tile == TILE_64 ? 4 : 3) * cus;
}

constexpr double kHbm = 2.7e12;       // bytes/s of the gathering sum / combine kernels (strided gathers)
constexpr double kHbmSums = 4.5e12;   // the one-pass sum kernels (translate sub-blocks, unit-stride K)
constexpr double kHbmComb = 4.0e12;   // the one-pass combine (every staged product read once)
constexpr double kLaunch = 4.0e-6;    // seconds per launch (dispatch and tail)

// `cpen`: epilogue_penalty of an ABC stage launch; `store`: the epilogue stores the product (AB, Naive).
double launch_s(int elem, int tile, int terms, bool dma, int64_t tiles, double flops, int64_t ks, double cpen,
                bool store)
{
    if (tiles <= 0 || flops <= 0) return 0.0;
    const int64_t slots = resident_slots(tile);
    const double waves = std::ceil(double(tiles) / double(slots)) * double(slots) / double(tiles);
    return flops / (leaf_rate(elem, tile, terms, dma, ks, store) * 1e12 * (1.0 - cpen)) * waves + kLaunch;
}

int64_t ntiles(int tile, int64_t ms, int64_t ns)
{
    const int64_t tm = tile_m(tile), tn = tile_n(tile);
    return ((ms + tm - 1) / tm) * ((ns + tn - 1) / tn);
}

}  // namespace

double model_ms(const Plan& p0)
{
    Plan& p = const_cast<Plan&>(p0);
    const int tile = p.tile;
    const int e = p.elem;
    double s = 0.0;
    const int64_t nb = p.nbatch;
    // single-term launches over K-fast operands (both K-fast, or staged copies) load straight into LDS
    const bool kfast = p.outer_staged || (!p.a_mfast && !p.b_nfast);
    // `per`: C targets per product of an ABC stage launch (0: no stage epilogue penalty)
    auto gemm = [&](int64_t ms, int64_t ns, int64_t ks, int nprod, int terms, int t, bool dma, double per,
                    bool store) {
        // rectangular tiles serve single-term launches with affine K only (else the launch takes 128 x 128)
        if (tile_is_rect(t) && terms > 1) t = TILE_128;
        // work of whole register tiles (a partly filled tile costs as much as a full one)
        const int64_t nt = ntiles(t, ms, ns) * nprod * nb;
        const double te = double(tile_m(t)) * double(tile_n(t));
        const double cpen = per > 0.0 ? epilogue_penalty(t, per, ks) : 0.0;
        return launch_s(e, t, terms, dma, nt, 2.0 * te * double(ks) * double(nt), ks, cpen, store);
    };
    // fringe passes (classical launches with 64 x 64 tiles; modelled with the plan's square tile)
    for (const Pass& f : p.fringe) s += gemm(f.ms, f.ns, f.ks, 1, 1, tile_is_rect(tile) ? TILE_64 : tile, kfast, 0.0, false);
    if (p.nprod == 0) return s * 1e3;
    if (!p.core)
    {
        s += gemm(p.ms, p.ns, p.ks, 1, 1, tile, kfast, 0.0, false);
        if (p.outer_staged)  // packed copies of A and B: read and written once each
            s += 2.0 * (double(p.ms) * p.ks + double(p.ks) * p.ns) * e / kHbm + 2 * kLaunch;
        return s * 1e3;
    }
    const double cbytes = double(p.ms) * p.ns * e * nb;
    // combine: read every M_p once and update every C sub-block (one-pass form), or per sub-block
    const double comb = p.h_comb_mask.empty() ? (double(p.h_terms.size()) + 2.0 * p.nsub) * cbytes / kHbm
                                              : (double(p.nprod) + 2.0 * p.nsub) * cbytes / kHbmComb;
    if (p.impl == IMPL_AB)
    {
        s += gemm(p.ms, p.ns, p.ks, p.nprod, max_terms(p.h_prods.data(), p.nprod), tile, false, 0.0, true);
        s += comb + kLaunch;
    }
    else if (p.impl == IMPL_NAIVE)
    {
        // sums: one pass over each operand when the tables allow it (S then K-fast, the dense GEMM loads
        // straight into LDS), else one gathering pass per sum
        const bool d2 = p.msum_nsrc == int(p.d * p.d);
        // (the dense launch's K check runs at upload; a host plan assumes it passes with K-fast sums)
        const bool one_a = d2 && p.msum_a && !p.a_mfast, one_b = d2 && p.msum_b && !p.b_nfast;
        const double sa = double(p.ms) * p.ks, sb = double(p.ks) * p.ns;
        double ta = 0, tb = 0;
        for (int q = 0; q < p.nprod; q++)
        {
            ta += p.h_prods[q].nA;
            tb += p.h_prods[q].nB;
        }
        s += (one_a ? (double(p.d * p.d) + p.nprod) * sa * e * nb / kHbmSums : (ta + p.nprod) * sa * e * nb / kHbm);
        s += (one_b ? (double(p.d * p.d) + p.nprod) * sb * e * nb / kHbmSums : (tb + p.nprod) * sb * e * nb / kHbm);
        s += 2 * kLaunch;
        s += gemm(p.ms, p.ns, p.ks, p.nprod, 1,
                  p.tile_given ? p.tile : (p.ms * p.ns >= int64_t(1024) * 1024 ? TILE_128 : TILE_64), one_a && one_b,
                  0.0, true);
        s += comb + kLaunch;
    }
    else
    {
        std::vector<int> order;
        make_stages(p, order);
        for (size_t st = 0; st + 1 < p.h_stage_begin.size(); st++)
        {
            int terms = p.h_stage_terms[st];
            if (p.outer_staged)
            {
                terms = 1;
                for (int j = p.h_stage_begin[st]; j < p.h_stage_begin[st + 1]; j++)
                    terms = std::max(terms, max_terms(&p.h_inner[size_t(order[size_t(j)])], 1));
            }
            int nct = 0;
            for (int j = p.h_stage_begin[st]; j < p.h_stage_begin[st + 1]; j++) nct += p.h_prods[size_t(order[size_t(j)])].nC;
            const double per = double(nct) / double(p.h_stage_begin[st + 1] - p.h_stage_begin[st]);
            s += gemm(p.ms, p.ns, p.ks, p.h_stage_begin[st + 1] - p.h_stage_begin[st], terms, tile, kfast && terms == 1,
                      per, false);
        }
        if (p.outer_staged)
        {
            // staged sums: read every term's sub-block (one-pass form: every source sub-block once) and
            // write the dense sums, for A and for B
            double ra = 0, rb = 0;
            for (const ProdDesc& sd : p.h_outer_sums)
            {
                ra += sd.nA + 1;
                rb += sd.nB + 1;
            }
            const double nout = double(p.h_outer_sums.size());
            if (p.msum_a && !p.a_mfast) ra = p.msum_nsrc + nout;
            if (p.msum_b && !p.b_nfast) rb = p.msum_nsrc + nout;
            // translate-form sums stream at about 5 TB/s, gathered ones at kHbm
            const double ha = p.outer_shift_a && !p.a_mfast ? kHbmSums : kHbm;
            const double hb = p.outer_shift_b && !p.b_nfast ? kHbmSums : kHbm;
            s += ra * double(p.os_m) * p.os_k * e / ha + rb * double(p.os_k) * p.os_n * e / hb + 2 * kLaunch;
        }
    }
    return s * 1e3;
}

// The classical launch, staged ABC stages, Naive's dense GEMM.
bool has_single_term_launch(const Plan& p)
{
    if (!p.core) return true;
    if (p.impl == IMPL_NAIVE) return true;
    if (p.impl == IMPL_ABC && p.outer_staged)
    {
        // the products over the staged sums (host plans have them before any device image is built)
        for (const ProdDesc& q : p.h_inner)
            if (q.nA > 1 || q.nB > 1) return false;
        return !p.h_inner.empty();
    }
    return false;
}

Status tune_candidates(const tblis::tblis_tensor* A, const label_type* idx_A, const tblis::tblis_tensor* B,
                       const label_type* idx_B, const tblis::tblis_tensor* C, const label_type* idx_C,
                       std::vector<TuneCand>& cands, bool& general)
{
    cands.clear();
    general = false;
    for (int level = 0; level <= 2; level++)
        for (int impl : {int(IMPL_ABC), int(IMPL_AB), int(IMPL_NAIVE)})
        {
            if (level == 0 && impl != IMPL_ABC) continue;
            for (int tile : {int(TILE_128), int(TILE_64), int(TILE_128x64), int(TILE_64x128)})
            {
                Plan p;
                Status st = build_plan_host(A, idx_A, B, idx_B, C, idx_C, level, impl, tile, 0, p);
                if (!st.ok()) return st;
                if (p.general)
                {
                    general = true;
                    return Status();
                }
                if (level > 0 && !p.core) continue;  // would only repeat the classical launch
                if (tile_is_rect(tile) && !has_single_term_launch(p)) continue;  // the same launches as 128 x 128
                cands.push_back(TuneCand{level, impl, tile, model_ms(p)});
            }
        }
    std::sort(cands.begin(), cands.end(), [](const TuneCand& a, const TuneCand& b) { return a.ms < b.ms; });
    return Status();
}

Status autotune(const tblis::tblis_tensor* A, const label_type* idx_A, const tblis::tblis_tensor* B,
                const label_type* idx_B, const tblis::tblis_tensor* C, const label_type* idx_C, bool measure,
                hipStream_t stream, TuneChoice& out)
{
    const std::string key = shape_key(A, idx_A, B, idx_B, C, idx_C);
    {
        std::lock_guard<std::mutex> lk(g_tune_mutex);
        auto it = g_measured.find(key);
        if (it != g_measured.end() && (it->second.measured || !measure))
        {
            out = it->second;
            return Status();
        }
    }
    std::vector<TuneCand> cands;
    bool general = false;
    Status st0 = tune_candidates(A, idx_A, B, idx_B, C, idx_C, cands, general);
    if (!st0.ok()) return st0;
    if (general)
    {
        // the general path reduces operands first; its inner contraction takes the default
        out = TuneChoice{1, IMPL_ABC, -1, 0.0, false};
        return Status();
    }
    TuneChoice best{cands[0].level, cands[0].impl, cands[0].tile, cands[0].ms, false};

    if (measure)
    {
        for (const void* ptr : {static_cast<const void*>(A->data), static_cast<const void*>(B->data),
                                static_cast<const void*>(C->data)})
        {
            hipPointerAttribute_t attr;
            if (!ptr || hipPointerGetAttributes(&attr, ptr) != hipSuccess || attr.type != hipMemoryTypeDevice)
            {
                (void)hipGetLastError();
                return Status::error(MISTRA_ERR_ARG, "measured tuning needs device tensors");
            }
        }
        // time every candidate on a scratch copy of C (same strides): one untimed call, then the minimum
        // over 3 rounds of 3 back-to-back calls (per call); calls timed one at a time, each between its
        // own events, carried the host's time to issue the call, several percent of a 1 ms call
        const int elem = C->type == tblis::TYPE_DOUBLE ? 8 : 4;
        int64_t lo = 0, hi = 0;
        for (unsigned i = 0; i < C->ndim; i++)
        {
            if (C->len[i] <= 0) return Status::error(MISTRA_ERR_ARG, "empty C");
            const int64_t ext = (C->len[i] - 1) * C->stride[i];
            (ext < 0 ? lo : hi) += ext;
        }
        const size_t bytes = size_t(hi - lo + 1) * elem;
        void* scratch = nullptr;
        hipError_t e = stream_alloc(&scratch, bytes, stream);
        if (e != hipSuccess) return Status::error(MISTRA_ERR_NOMEM, hipGetErrorString(e));
        e = hipMemcpyAsync(scratch, static_cast<const char*>(C->data) + lo * elem, bytes, hipMemcpyDeviceToDevice,
                           stream);
        hipEvent_t t0, t1;
        (void)hipEventCreate(&t0);
        (void)hipEventCreate(&t1);
        tblis::tblis_tensor Cs = *C;
        Cs.data = static_cast<char*>(scratch) - lo * elem;
        double best_ms = 1e30;
        Status st;
        for (const TuneCand& c : cands)
        {
            double ms_min = 1e30;
            for (int r = 0; r < 4 && st.ok(); r++)
            {
                const int reps = r == 0 ? 1 : 3;
                (void)hipEventRecord(t0, stream);
                for (int i = 0; i < reps && st.ok(); i++)
                {
                    tblis::tblis_tensor Ct = Cs;  // tensor_mult resets the scalar of C after each call
                    st = tensor_mult(A, idx_A, B, idx_B, &Ct, idx_C, c.level, c.impl, MISTRA_PREC_NATIVE, 0, stream,
                                     false, c.tile);
                }
                (void)hipEventRecord(t1, stream);
                if (!st.ok()) break;
                (void)hipEventSynchronize(t1);
                float ms = 0;
                (void)hipEventElapsedTime(&ms, t0, t1);
                if (r > 0) ms_min = std::min(ms_min, double(ms) / reps);
            }
            if (!st.ok()) break;
            if (ms_min < best_ms)
            {
                best_ms = ms_min;
                best = TuneChoice{c.level, c.impl, c.tile, ms_min, true};
            }
        }
        (void)hipEventDestroy(t0);
        (void)hipEventDestroy(t1);
        (void)hipFreeAsync(scratch, stream);
        (void)hipStreamSynchronize(stream);
        if (!st.ok()) return st;
        if (e != hipSuccess) return Status::error(MISTRA_ERR_HIP, hipGetErrorString(e));
    }
    {
        std::lock_guard<std::mutex> lk(g_tune_mutex);
        if (g_measured.size() >= kMaxTuneEntries) g_measured.clear();
        g_measured[key] = best;
    }
    out = best;
    return Status();
}

int best_level(const tblis::tblis_tensor* A, const label_type* idx_A, const tblis::tblis_tensor* B,
               const label_type* idx_B, const tblis::tblis_tensor* C, const label_type* idx_C, int impl)
{
    // cached per shape: batches and loops ask for the same shapes on every call
    const std::string key = shape_key(A, idx_A, B, idx_B, C, idx_C) + '#' + std::to_string(impl);
    {
        std::lock_guard<std::mutex> lk(g_tune_mutex);
        auto it = g_best_level.find(key);
        if (it != g_best_level.end()) return it->second;
    }
    int best = 0;
    double best_ms = 1e300;
    for (int level = 0; level <= 2; level++)
    {
        Plan p;
        if (!build_plan_host(A, idx_A, B, idx_B, C, idx_C, level, impl, -1, 0, p).ok() || p.general) return 1;
        if (level > 0 && !p.core) continue;
        const double ms = model_ms(p);
        if (ms < best_ms)
        {
            best_ms = ms;
            best = level;
        }
    }
    std::lock_guard<std::mutex> lk(g_tune_mutex);
    if (g_best_level.size() >= kMaxTuneEntries) g_best_level.clear();
    g_best_level[key] = best;
    return best;
}

}  // namespace mistra
