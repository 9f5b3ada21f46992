// Host driver: plan construction and caching, launch sequencing per level and impl.
//
// Reference call structure being replaced (src/stra_internal/3t/):
//   stra_mult dispatcher          stra_mult.cxx:1701-1846, stra_mult_2level.cxx:4434-4578
//   ABC, 7 products serialized    stra_mult.cxx:308-605 (L=1), stra_mult_2level.cxx:241-1486 (L=2)
//   Naive (explicit S, T, M)      stra_mult.cxx:630-951, straprim_common.hpp:70-157
//   AB (staged M, scatter-add)    stra_mult.cxx:976-1285, straprim_common.hpp:160-204
// Here a level/impl becomes: one ABC launch that runs every product per tile; or, for AB, one launch
// computing all products into HBM staging plus one combine pass; or, for Naive, two operand-sum
// passes, one dense GEMM launch and one combine pass.  The rows/cols/k that the 2^L split leaves over
// (dropped by the reference, src/matrix/stra_tensor_view.hpp:246-247) are computed by classical
// passes unless MISTRA_FLAG_REF_FRINGE asks for the reference's truncation.
#include "driver.hpp"
#include "lowp.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <list>
#include <map>
#include <mutex>
#include <sstream>
#include <tuple>
#include <unordered_map>

namespace mistra {

namespace {

std::mutex g_plan_mutex;
using PlanList = std::list<std::pair<std::string, std::shared_ptr<Plan>>>;
PlanList g_plans;                                                  // LRU, most recent first
std::unordered_map<std::string, PlanList::iterator> g_plan_index;  // key -> its entry of g_plans
constexpr size_t kMaxPlans = 1024;  // plans are small: index vectors and tables
// Plans whose device tables a caller's HIP graph has captured: a graph keeps the table addresses in its
// kernel nodes and may be replayed at any later time, so these plans stay resident until the process
// ends (neither LRU eviction nor mistra_release_workspace frees them).
std::unordered_map<std::string, std::shared_ptr<Plan>> g_pinned;

#define HIP_TRY(expr)                                                                             \
    do                                                                                            \
    {                                                                                             \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return Status::error(MISTRA_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

void append_key(std::string& key, const tblis::tblis_tensor* t, const label_type* idx)
{
    std::ostringstream os;
    os << t->ndim << ':';
    for (unsigned i = 0; i < t->ndim; i++) os << idx[i] << t->len[i] << ',' << t->stride[i] << ';';
    key += os.str();
}

bool unit(const std::vector<int64_t>& v)
{
    return v.size() > 1 && (v[1] - v[0] == 1 || v[1] - v[0] == -1);
}

template <typename V>
void permute_group(IndexGroup& g, const std::vector<unsigned>& order)
{
    IndexGroup out;
    for (unsigned i : order)
    {
        out.idx.push_back(g.idx[i]);
        out.len.push_back(g.len[i]);
        if (!g.sA.empty()) out.sA.push_back(g.sA[i]);
        if (!g.sB.empty()) out.sB.push_back(g.sB[i]);
        if (!g.sC.empty()) out.sC.push_back(g.sC[i]);
    }
    g = std::move(out);
}

ProdDesc make_desc(const Product& pr, int64_t d, int64_t ms, int64_t ns, int64_t ks, unsigned& touched)
{
    ProdDesc pd;
    std::memset(&pd, 0, sizeof(pd));
    pd.nA = int(pr.A.size());
    pd.nB = int(pr.B.size());
    pd.nC = int(pr.C.size());
    for (size_t i = 0; i < pr.A.size(); i++)
        pd.a[i] = SubRef{0, (pr.A[i].subid / d) * ms, (pr.A[i].subid % d) * ks, pr.A[i].coeff};
    for (size_t i = 0; i < pr.B.size(); i++)
        pd.b[i] = SubRef{0, (pr.B[i].subid / d) * ks, (pr.B[i].subid % d) * ns, pr.B[i].coeff};
    for (size_t i = 0; i < pr.C.size(); i++)
    {
        pd.c[i] = SubRef{0, (pr.C[i].subid / d) * ms, (pr.C[i].subid % d) * ns, pr.C[i].coeff};
        if (!(touched & (1u << pr.C[i].subid)))
        {
            pd.first_mask |= 1 << int(i);
            touched |= 1u << pr.C[i].subid;
        }
    }
    return pd;
}

ProdDesc single_desc(int64_t a_r, int64_t a_c, int64_t b_r, int64_t b_c, int64_t c_r, int64_t c_c, bool first)
{
    ProdDesc pd;
    std::memset(&pd, 0, sizeof(pd));
    pd.nA = pd.nB = pd.nC = 1;
    pd.a[0] = SubRef{0, a_r, a_c, 1.0};
    pd.b[0] = SubRef{0, b_r, b_c, 1.0};
    pd.c[0] = SubRef{0, c_r, c_c, 1.0};
    pd.first_mask = first ? 1 : 0;
    return pd;
}

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// Workgroup tile of the leaf kernel, per launch kind.  MISTRA_TILE=64|128 overrides (for tuning).
// Measured on MI355X at m=n=k=8192 (FP64): launches with the multi-C epilogue (classical and ABC) and
// the AB staging launch run best with 64x64 tiles (fewer registers: 3-4 waves per SIMD); the dense
// Naive GEMM over contiguous S_p / T_p with 128x128 tiles.
int forced_tile()
{
    static const int forced = [] {
        const char* e = std::getenv("MISTRA_TILE");
        if (!e) return -1;
        const std::string v(e);
        return v == "64" ? int(TILE_64) : v == "128" ? int(TILE_128) : v == "128x64" ? int(TILE_128x64)
                                                           : v == "64x128" ? int(TILE_64x128) : -1;
    }();
    return forced;
}

// 128 x 128 register tiles once a leaf has enough of them to fill the GPU (2 per CU at 256 CUs);
// smaller leaves take 64 x 64 tiles (4 per CU).
int choose_tile(const Plan& p)
{
    if (forced_tile() >= 0) return forced_tile();
    const int64_t ms = p.core ? p.ms : p.m, ns = p.core ? p.ns : p.n;
    const int64_t t128 = ((ms + 127) / 128) * ((ns + 127) / 128) * p.nbatch;
    return t128 >= 256 ? TILE_128 : TILE_64;
}

int dense_tile(const Plan& p)
{
    if (forced_tile() >= 0) return forced_tile();
    if (p.tile_given) return p.tile;
    return p.ms * p.ns >= int64_t(1024) * 1024 ? TILE_128 : TILE_64;
}

// Whether the sub-block at `off` (length len) of scatter vector v is a translate of the one at 0:
// v[off + i] - v[off] == v[i] - v[0] for every i < len.
bool is_translate(const std::vector<int64_t>& v, int64_t off, int64_t len)
{
    if (off == 0) return true;
    const int64_t d = v[size_t(off)] - v[0];
    for (int64_t i = 0; i < len; i++)
        if (v[size_t(off + i)] - v[size_t(i)] != d) return false;
    return true;
}

// d if v[k] = v[0] + k d for every k < len (len >= 1), else -1 (also for a negative d)
int64_t affine_step(const std::vector<int64_t>& v, int64_t len)
{
    if (len < 1 || int64_t(v.size()) < len) return -1;
    const int64_t d = len > 1 ? v[1] - v[0] : 0;
    if (d < 0) return -1;
    for (int64_t k = 0; k < len; k++)
        if (v[size_t(k)] != v[0] + k * d) return -1;
    return d;
}

// A launch may use the fast kernel body when every sum term's sub-block (A, B and the C targets) is a
// translate of the first sub-block in its operand's scatter vectors, so that a term is the first
// sub-block shifted by a uniform offset (the fast product table stores it as the term's base).  Then:
// KMODE_AFFINE if the first sub-block's K runs are affine at every K-tile start with one stride per
// operand and the offsets inside a K-tile fit 31 bits; else KMODE_GATHER if row + K offsets fit 31 bits;
// else (and without the translate property) the general body.
KFast kfast_check(const Plan& p, const std::vector<ProdDesc>& prods, int first, int count, int64_t ms, int64_t ns,
                  int64_t ks, int TK, const std::vector<int64_t>& a_krun, const std::vector<int64_t>& b_krun)
{
    KFast kf;
    if (ks <= 0 || ms <= 0 || ns <= 0) return kf;
    for (int q = first; q < first + count; q++)
    {
        const ProdDesc& pd = prods[size_t(q)];
        for (int t = 0; t < pd.nA; t++)
            if (!is_translate(p.h_a_m, pd.a[t].roff, ms) || !is_translate(p.h_a_k, pd.a[t].coff, ks)) return kf;
        for (int t = 0; t < pd.nB; t++)
            if (!is_translate(p.h_b_k, pd.b[t].roff, ks) || !is_translate(p.h_b_n, pd.b[t].coff, ns)) return kf;
        for (int t = 0; t < pd.nC; t++)
            if (!is_translate(p.h_c_m, pd.c[t].roff, ms) || !is_translate(p.h_c_n, pd.c[t].coff, ns)) return kf;
    }
    const int64_t limit = (int64_t(1) << 31) / p.elem;
    int64_t am = 0, bn = 0;
    for (int64_t i = 0; i < ms; i++)
    {
        if (p.h_a_m[size_t(i)] < 0) return kf;
        am = std::max(am, p.h_a_m[size_t(i)]);
    }
    for (int64_t i = 0; i < ns; i++)
    {
        if (p.h_b_n[size_t(i)] < 0) return kf;
        bn = std::max(bn, p.h_b_n[size_t(i)]);
    }
    // affine K runs with one stride per operand at every K-tile start
    const int64_t sa = a_krun[0], sb = b_krun[0];
    bool affine = sa > 0 && sb > 0;
    for (int64_t k0 = 0; affine && k0 < ks; k0 += TK)
        affine = a_krun[size_t(k0)] == sa && b_krun[size_t(k0)] == sb;
    if (affine && am + (TK - 1) * sa < limit && bn + (TK - 1) * sb < limit)
    {
        kf.on = true;
        // whole K-tiles and a K-fast operand: 16-byte loads along its unit-stride K
        kf.mode = ks % TK == 0 && (!p.a_mfast || !p.b_nfast) ? KMODE_AFFINE_VEC : KMODE_AFFINE;
        kf.sa = sa;
        kf.sb = sb;
        return kf;
    }
    // gathered K offsets: row offset + K offset must fit the 31-bit buffer range
    int64_t ak = 0, bk = 0;
    for (int64_t k = 0; k < ks; k++)
    {
        if (p.h_a_k[size_t(k)] < 0 || p.h_b_k[size_t(k)] < 0) return kf;
        ak = std::max(ak, p.h_a_k[size_t(k)]);
        bk = std::max(bk, p.h_b_k[size_t(k)]);
    }
    // (and the kernel's 32-bit byte offsets into the K tables: ks < 2^28)
    if (am + ak < limit && bn + bk < limit && ks < (int64_t(1) << 28))
    {
        kf.on = true;
        kf.mode = KMODE_GATHER;
    }
    return kf;
}

// The fast body's product table: each term's base becomes its uniform shift from the first sub-block.
ProdDesc shifted(const Plan& p, const ProdDesc& pd)
{
    auto d = [](const std::vector<int64_t>& v, int64_t off) { return v[size_t(off)] - v[0]; };
    ProdDesc f = pd;
    for (int t = 0; t < pd.nA; t++) f.a[t].base = pd.a[t].base + d(p.h_a_m, pd.a[t].roff) + d(p.h_a_k, pd.a[t].coff);
    for (int t = 0; t < pd.nB; t++) f.b[t].base = pd.b[t].base + d(p.h_b_k, pd.b[t].roff) + d(p.h_b_n, pd.b[t].coff);
    for (int t = 0; t < pd.nC; t++) f.c[t].base = pd.c[t].base + d(p.h_c_m, pd.c[t].roff) + d(p.h_c_n, pd.c[t].coff);
    return f;
}

// One-pass sum tables (launch_operand_multisums) for the sums of the first `depth` levels over sub-blocks
// of extent om x ok (A) and ok x on (B): source s = sub-block s (row-major) of the 2^depth split; A's
// rows (K) are its column blocks, B's rows (K) its row blocks.  msum_a / msum_b say whether every source
// sub-block is a translate of the first.
void build_msum(Plan& p, int depth, int64_t om, int64_t on, int64_t ok)
{
    const int64_t dout = int64_t(1) << depth;
    const auto& one = products_for_level(depth);
    const int nsrc = int(dout * dout);
    p.msum_nsrc = nsrc;
    p.h_msum_shift_a.assign(size_t(nsrc), 0);
    p.h_msum_shift_b.assign(size_t(nsrc), 0);
    p.msum_a = p.msum_b = true;
    for (int src = 0; src < nsrc; src++)
    {
        const int64_t rb = src / dout, cb = src % dout;
        p.msum_a = p.msum_a && is_translate(p.h_a_k, cb * ok, ok) && is_translate(p.h_a_m, rb * om, om);
        p.msum_b = p.msum_b && is_translate(p.h_b_k, rb * ok, ok) && is_translate(p.h_b_n, cb * on, on);
        if (p.msum_a)
            p.h_msum_shift_a[size_t(src)] =
                (p.h_a_k[size_t(cb * ok)] - p.h_a_k[0]) + (p.h_a_m[size_t(rb * om)] - p.h_a_m[0]);
        if (p.msum_b)
            p.h_msum_shift_b[size_t(src)] =
                (p.h_b_k[size_t(rb * ok)] - p.h_b_k[0]) + (p.h_b_n[size_t(cb * on)] - p.h_b_n[0]);
    }
    auto pairs = [](const std::vector<int64_t>& rs, int64_t rows, const std::vector<int64_t>& cs, int64_t cols,
                    const std::vector<int64_t>& shift) {
        if (rows % 2 != 0) return false;
        for (int64_t x : shift)
            if (x % 2 != 0) return false;
        for (int64_t r = 0; r + 1 < rows; r += 2)
            if (rs[size_t(r)] % 2 != 0 || rs[size_t(r + 1)] != rs[size_t(r)] + 1) return false;
        for (int64_t c = 0; c < cols; c++)
            if (cs[size_t(c)] % 2 != 0) return false;
        return true;
    };
    p.msum_pair_a = p.msum_a && pairs(p.h_a_k, ok, p.h_a_m, om, p.h_msum_shift_a);
    p.msum_pair_b = p.msum_b && pairs(p.h_b_k, ok, p.h_b_n, on, p.h_msum_shift_b);
    p.h_msum_coef_a.assign(one.size() * size_t(nsrc), 0.0);
    p.h_msum_coef_b.assign(one.size() * size_t(nsrc), 0.0);
    for (size_t o = 0; o < one.size(); o++)
    {
        for (const SubTerm& t : one[o].A) p.h_msum_coef_a[o * size_t(nsrc) + size_t(t.subid)] = t.coeff;
        for (const SubTerm& t : one[o].B) p.h_msum_coef_b[o * size_t(nsrc) + size_t(t.subid)] = t.coeff;
    }
    // the 2-level coefficients are the Kronecker square of the 1-level ones (plan.cpp: kron)
    const auto& l1 = strassen_products_1level();
    p.h_msum_coef1_a.assign(depth == 2 ? 28 : 0, 0.0);
    p.h_msum_coef1_b.assign(depth == 2 ? 28 : 0, 0.0);
    for (size_t o = 0; depth == 2 && o < l1.size(); o++)
    {
        for (const SubTerm& t : l1[o].A) p.h_msum_coef1_a[o * 4 + size_t(t.subid)] = t.coeff;
        for (const SubTerm& t : l1[o].B) p.h_msum_coef1_b[o * 4 + size_t(t.subid)] = t.coeff;
    }
}

// Tables of the staged form of an ABC plan (see Plan::outer_staged), staging the sums of the first
// `depth` levels.  Outer product po of the depth-level table sums A over its terms into S_po (element
// (i, j) of the os_m x os_k block at po*os_m*os_k + i*os_k + j: K-fast) and B into T_po (element (i, j)
// of os_k x os_n at po*os_k*os_n + i + j*os_k: K-fast); product n_in*po + pi of the plan's table then
// multiplies the sums of inner product pi (of the (level-depth)-level table: the identity when every
// level is staged) over S_po and T_po and updates the same C targets, with the same coefficients, as
// before.  Returns false when the form does not apply (C sub-blocks that are not translates of the
// first one, or an inner sub-block that does not fit the 31-bit buffer range).
bool build_outer_staged(Plan& p, int depth)
{
    const int64_t dout = int64_t(1) << depth;  // staged split divisor
    const int64_t f = p.d / dout;              // leaf blocks per staged sub-block side
    const int64_t ms = p.ms, ns = p.ns, ks = p.ks, om = f * ms, on = f * ns, ok = f * ks;
    for (int q = 0; q < p.nprod; q++)
    {
        const ProdDesc& pd = p.h_prods[size_t(q)];
        for (int t = 0; t < pd.nC; t++)
            if (!is_translate(p.h_c_m, pd.c[t].roff, ms) || !is_translate(p.h_c_n, pd.c[t].coff, ns)) return false;
    }
    const int TK = tile_kdepth(p.tile, p.elem);
    const int64_t limit = (int64_t(1) << 31) / p.elem;
    if ((ms - 1) * ok + ks + TK >= limit || (ns - 1) * ok + ks + TK >= limit) return false;
    const auto& one = products_for_level(depth);
    p.staged_depth = depth;
    p.os_m = om;
    p.os_n = on;
    p.os_k = ok;
    p.h_outer_sums.clear();
    for (const Product& po : one)
    {
        ProdDesc sd;
        std::memset(&sd, 0, sizeof(sd));
        sd.nA = int(po.A.size());
        sd.nB = int(po.B.size());
        // A: rows are K (a_k), columns M (a_m); B: rows K (b_k), columns N (b_n)
        for (size_t t = 0; t < po.A.size(); t++)
            sd.a[t] = SubRef{0, (po.A[t].subid % dout) * ok, (po.A[t].subid / dout) * om, po.A[t].coeff};
        for (size_t t = 0; t < po.B.size(); t++)
            sd.b[t] = SubRef{0, (po.B[t].subid / dout) * ok, (po.B[t].subid % dout) * on, po.B[t].coeff};
        p.h_outer_sums.push_back(sd);
    }
    // translate form of the outer sums (rows along K): shifts per term, and 2-row vector reads when
    // K comes in contiguous pairs at even offsets
    auto translate_form = [&](int which, const std::vector<int64_t>& rs, const std::vector<int64_t>& cs,
                              int64_t rows, int64_t cols, int64_t zoff, bool& shift, bool& pair) {
        shift = true;
        bool even = zoff % 2 == 0 && rows % 2 == 0;
        for (size_t q = 0; q < p.h_outer_sums.size(); q++)
        {
            const ProdDesc& sd = p.h_outer_sums[q];
            const int nt = which == 0 ? sd.nA : sd.nB;
            const SubRef* t = which == 0 ? sd.a : sd.b;
            SubRef* f = which == 0 ? p.h_outer_shifted[q].a : p.h_outer_shifted[q].b;
            for (int i = 0; i < nt; i++)
            {
                shift = shift && is_translate(rs, t[i].roff, rows) && is_translate(cs, t[i].coff, cols);
                if (!shift) return;
                f[i].base = (rs[size_t(t[i].roff)] - rs[0]) + (cs[size_t(t[i].coff)] - cs[0]);
                even = even && f[i].base % 2 == 0;
            }
        }
        for (int64_t r = 0; even && r + 1 < rows; r += 2) even = rs[size_t(r)] % 2 == 0 && rs[size_t(r + 1)] == rs[size_t(r)] + 1;
        for (int64_t c = 0; even && c < cols; c++) even = cs[size_t(c)] % 2 == 0;
        pair = even;
    };
    p.h_outer_shifted = p.h_outer_sums;
    translate_form(0, p.h_a_k, p.h_a_m, ok, om, p.h_batch_a[0], p.outer_shift_a, p.outer_pair_a);
    translate_form(1, p.h_b_k, p.h_b_n, ok, on, p.h_batch_b[0], p.outer_shift_b, p.outer_pair_b);
    build_msum(p, depth, om, on, ok);
    p.h_inner.clear();
    const auto& inner = products_for_level(p.level - depth);
    const int64_t din = f;  // inner split divisor
    for (size_t o = 0; o < one.size(); o++)
        for (size_t i = 0; i < inner.size(); i++)
        {
            const Product& pi = inner[i];
            ProdDesc pd = p.h_prods[o * inner.size() + i];  // C terms of the plan's product
            pd.nA = int(pi.A.size());
            pd.nB = int(pi.B.size());
            std::memset(pd.a, 0, sizeof(pd.a));
            std::memset(pd.b, 0, sizeof(pd.b));
            // fast-body form: base = offset of the inner sub-block's first element
            for (size_t t = 0; t < pi.A.size(); t++)
            {
                const int64_t r = (pi.A[t].subid / din) * ms, c = (pi.A[t].subid % din) * ks;
                pd.a[t] = SubRef{int64_t(o) * om * ok + r * ok + c, r, c, pi.A[t].coeff};
            }
            for (size_t t = 0; t < pi.B.size(); t++)
            {
                const int64_t r = (pi.B[t].subid / din) * ks, c = (pi.B[t].subid % din) * ns;
                pd.b[t] = SubRef{int64_t(o) * ok * on + r + c * ok, r, c, pi.B[t].coeff};
            }
            auto d = [](const std::vector<int64_t>& v, int64_t off) { return v[size_t(off)] - v[0]; };
            for (int t = 0; t < pd.nC; t++) pd.c[t].base = d(p.h_c_m, pd.c[t].roff) + d(p.h_c_n, pd.c[t].coff);
            p.h_inner.push_back(pd);
        }
    p.kf_inner.on = true;
    p.kf_inner.mode = ks % TK == 0 ? KMODE_AFFINE_VEC : KMODE_AFFINE;
    p.kf_inner.sa = 1;
    p.kf_inner.sb = 1;
    return true;
}

// Whether (and how deep) a plan stages its operand sums: MISTRA_FLAG_FUSED_SUMS, MISTRA_FLAG_OUTER_STAGED
// (the outer level; at level 0 the packed copies) and MISTRA_FLAG_ALL_STAGED (every level) decide when
// given, else the cost model (the staged sums' HBM traffic against the slower multi-term or
// register-staged leaf launches), for leaves of at least one register tile per CU.
void choose_outer_staged(Plan& p, unsigned flags)
{
    p.outer_staged = false;
    p.staged_depth = 0;
    if (p.nprod == 0 || p.general || p.nbatch != 1 || (flags & MISTRA_FLAG_FUSED_SUMS)) return;
    // Level 0: staging at depth 0 packs A and B into K-fast dense copies (for the direct-to-LDS leaf
    // path), worth it only when an operand is not K-fast.  Levels 1, 2: ABC only.
    int lo = 1, hi = p.level;
    if (p.level == 0)
    {
        if (!p.a_mfast && !p.b_nfast) return;
        lo = hi = 0;
    }
    else if (!p.core || p.impl != IMPL_ABC)
        return;
    auto set = [&](int depth) {
        p.outer_staged = depth >= 0 && build_outer_staged(p, depth);
        if (!p.outer_staged) p.staged_depth = 0;
    };
    if (flags & (MISTRA_FLAG_OUTER_STAGED | MISTRA_FLAG_ALL_STAGED))
    {
        set(p.level == 0 ? 0 : (flags & MISTRA_FLAG_ALL_STAGED ? p.level : 1));
        return;
    }
    // small leaves: launches that do not fill the 256 CUs are latency-bound, so fewer sum terms gain
    // little against the extra sum launches and the workspace
    const int64_t tm = tile_m(p.tile), tn = tile_n(p.tile);
    if (((p.ms + tm - 1) / tm) * ((p.ns + tn - 1) / tn) < 256) return;
    int best = -1;
    double best_ms = model_ms(p);
    for (int depth = lo; depth <= hi; depth++)
    {
        set(depth);
        if (!p.outer_staged) break;
        const double ms = model_ms(p);
        if (ms < best_ms)
        {
            best_ms = ms;
            best = depth;
        }
    }
    set(best);
}

// CombineArgs::pmask of the plan's combine table, or empty when the table does not fit that form
void build_comb_mask(Plan& p)
{
    p.h_comb_mask.assign(size_t(p.nprod), 0u);
    const bool fits = (p.nsub == 1 || p.nsub == 4 || p.nsub == 16) && p.nprod >= 1 && p.nprod <= kCombineMaxProds;
    for (int s = 0; fits && s < p.nsub; s++)
        for (int t = p.h_term_begin[size_t(s)]; t < p.h_term_begin[size_t(s) + 1]; t++)
        {
            const CombineTerm& ct = p.h_terms[size_t(t)];
            unsigned& m = p.h_comb_mask[size_t(ct.prod)];
            if ((m >> s) & 1u || (ct.coeff != 1.0 && ct.coeff != -1.0))
            {
                p.h_comb_mask.clear();
                return;
            }
            m |= (1u << s) | (ct.coeff < 0 ? 1u << (16 + s) : 0u);
        }
    if (!fits) p.h_comb_mask.clear();
}

// MultiSumArgs::cv: the 1-level factor when the sums go through it, else the whole matrix if it fits
void set_msum_cv(const Plan& p, bool isA, MultiSumArgs& ma)
{
    const std::vector<double>& c1 = isA ? p.h_msum_coef1_a : p.h_msum_coef1_b;
    const std::vector<double>& c = isA ? p.h_msum_coef_a : p.h_msum_coef_b;
    const std::vector<double>& src = c1.empty() ? c : c1;
    ma.ncv = src.size() <= size_t(kMsumMaxCv) ? int(src.size()) : 0;
    for (int i = 0; i < ma.ncv; i++) ma.cv[i] = src[size_t(i)];
}

void set_comb_mask(const Plan& p, CombineArgs& ca)
{
    ca.dense = p.h_comb_mask.empty() || g_combine_per_sub.load() ? 0 : 1;
    for (size_t q = 0; q < p.h_comb_mask.size(); q++) ca.pmask[q] = p.h_comb_mask[q];
}

}  // namespace

std::atomic<int> g_combine_per_sub{0};
std::atomic<int> g_dbg_ozaki_slices{0}, g_dbg_i8_slices{0}, g_dbg_i8_fold{0};
std::atomic<int> g_dbg_tile{-1};
std::atomic<int> g_dbg_unfused_lowp{0};

Plan::~Plan()
{
    if (dmem) (void)hipFree(dmem);
}

int tile_override() { return forced_tile(); }

Status build_plan_host(const tblis::tblis_tensor* A, const label_type* idx_A, const tblis::tblis_tensor* B,
                       const label_type* idx_B, const tblis::tblis_tensor* C, const label_type* idx_C,
                       int level, int impl, int tile, unsigned flags, Plan& p)
{
    if (level < 0 || level > 2) return Status::error(MISTRA_ERR_ARG, "level must be 0, 1 or 2");
    if (A->type != B->type || A->type != C->type)
        return Status::error(MISTRA_ERR_ARG, "A, B and C must have the same type");
    if (A->type != tblis::TYPE_DOUBLE && A->type != tblis::TYPE_FLOAT)
        return Status::error(MISTRA_ERR_UNSUPPORTED, "only TYPE_FLOAT and TYPE_DOUBLE are supported");
    if (A->conj || B->conj || C->conj)
        return Status::error(MISTRA_ERR_UNSUPPORTED, "conjugation is not supported for real types");

    p.elem = A->type == tblis::TYPE_DOUBLE ? 8 : 4;
    p.level = level;
    p.impl = impl;
    p.tile = tile;
    p.tile_given = tile >= 0;
    p.flags = flags;

    Grouping g;
    std::string err;
    if (!group_indices(*A, idx_A, *B, idx_B, *C, idx_C, g, err)) return Status::error(MISTRA_ERR_ARG, err);
    p.grouping = g;
    if (!g.pure())
    {
        p.general = true;
        return Status();
    }

    // src/stra_internal/3t/stra_mult.cxx:336-356: AC by (C, A) strides, BC by (C, B), AB by (A, B).
    // M and N orders: the reference sorts by (C, operand) strides.  C's unit-stride index lies in one
    // of the two groups; the other group is sorted by its operand's strides instead, so that the operand
    // (not C, which is written along the first group) gets unit-stride rows.
    {
        auto has_unit = [](const std::vector<stride_type>& st, const std::vector<len_type>& len) {
            for (size_t i = 0; i < st.size(); i++)
                if (st[i] == 1 && len[i] > 1) return true;
            return false;
        };
        const bool c_in_ac = has_unit(g.AC.sC, g.AC.len), c_in_bc = has_unit(g.BC.sC, g.BC.len);
        permute_group<int>(g.AC, !c_in_ac && c_in_bc ? sort_by_stride(g.AC.sA, g.AC.sC)
                                                     : sort_by_stride(g.AC.sC, g.AC.sA));
        permute_group<int>(g.BC, !c_in_bc && c_in_ac ? sort_by_stride(g.BC.sB, g.BC.sC)
                                                     : sort_by_stride(g.BC.sC, g.BC.sB));
    }
    // K order: the reference sorts by (A, B) strides; a different order only changes the summation
    // order.  When no K index of A has unit stride but one of B does, sort by (B, A) instead, so that
    // the K runs of B (which become the rows of A when C is transposed) are unit-stride and load
    // coalesced.
    {
        bool a_unit = false, b_unit = false;
        for (size_t i = 0; i < g.AB.len.size(); i++)
        {
            a_unit = a_unit || (g.AB.sA[i] == 1 && g.AB.len[i] > 1);
            b_unit = b_unit || (g.AB.sB[i] == 1 && g.AB.len[i] > 1);
        }
        if (!a_unit && b_unit)
            permute_group<int>(g.AB, sort_by_stride(g.AB.sB, g.AB.sA));
        else
            permute_group<int>(g.AB, sort_by_stride(g.AB.sA, g.AB.sB));
    }

    std::vector<int64_t> a_m = scatter(g.AC.len, g.AC.sA), c_m = scatter(g.AC.len, g.AC.sC);
    std::vector<int64_t> a_k = scatter(g.AB.len, g.AB.sA), b_k = scatter(g.AB.len, g.AB.sB);
    std::vector<int64_t> b_n = scatter(g.BC.len, g.BC.sB), c_n = scatter(g.BC.len, g.BC.sC);
    p.h_batch_a = scatter(g.ABC.len, g.ABC.sA);
    p.h_batch_b = scatter(g.ABC.len, g.ABC.sB);
    p.h_batch_c = scatter(g.ABC.len, g.ABC.sC);
    p.nbatch = int64_t(p.h_batch_c.size());

    // If C's column index is the unit-stride one, compute C^T = B^T A^T so that the MFMA output
    // (4 consecutive rows per lane) lands on contiguous memory (src/stra_internal/3t/stra_mult.cxx:436-444).
    if (unit(c_n) && !unit(c_m))
    {
        p.transposed = true;
        std::swap(a_m, b_n);  // A' = B^T : rows = n, cols = k
        std::swap(a_k, b_k);  // B' = A^T : rows = k, cols = m
        std::swap(c_m, c_n);
        std::swap(p.h_batch_a, p.h_batch_b);
    }
    p.m = int64_t(a_m.size());
    p.n = int64_t(b_n.size());
    p.k = int64_t(a_k.size());
    if (p.k / 16 >= int64_t(INT32_MAX))
        return Status::error(MISTRA_ERR_UNSUPPORTED, "contracted extent of 2^35 or more (K-tiles are counted in 32 bits)");
    p.a_mfast = unit(a_m) || !unit(a_k);
    p.b_nfast = unit(b_n) || !unit(b_k);

    p.d = int64_t(1) << level;
    p.ms = p.m / p.d;
    p.ns = p.n / p.d;
    p.ks = p.k / p.d;
    p.core = level > 0 && p.ms > 0 && p.ns > 0 && p.ks > 0;
    if (p.tile < 0) p.tile = choose_tile(p);
    if (level > 0 && !p.core && (flags & MISTRA_FLAG_REF_FRINGE))
    {
        // the reference computes nothing when a leaf extent truncates to zero
        p.h_prods.clear();
        p.nprod = 0;
    }
    else if (!p.core)
    {
        p.d = 1;
        p.ms = p.m;
        p.ns = p.n;
        p.ks = p.k;
        p.h_prods = {single_desc(0, 0, 0, 0, 0, 0, true)};
        p.nprod = 1;
        // one C block fed by the one product (used when the product is staged, as in the low-precision path)
        p.nsub = 1;
        p.h_term_begin = {0, 1};
        p.h_sub_roff = {0};
        p.h_sub_coff = {0};
        p.h_terms = {CombineTerm{0, 1.0}};
        build_comb_mask(p);
    }
    else
    {
        unsigned touched = 0;
        for (const Product& pr : products_for_level(level))
            p.h_prods.push_back(make_desc(pr, p.d, p.ms, p.ns, p.ks, touched));
        p.nprod = int(p.h_prods.size());
        if (!(flags & MISTRA_FLAG_REF_FRINGE))
        {
            const int64_t dm = p.d * p.ms, dn = p.d * p.ns, dk = p.d * p.ks;
            if (dm < p.m)
            {
                p.fringe.push_back(Pass{p.m - dm, p.n, p.k, int(p.h_prods.size())});
                p.h_prods.push_back(single_desc(dm, 0, 0, 0, dm, 0, true));
            }
            if (dn < p.n)
            {
                p.fringe.push_back(Pass{dm, p.n - dn, p.k, int(p.h_prods.size())});
                p.h_prods.push_back(single_desc(0, 0, 0, dn, 0, dn, true));
            }
            if (dk < p.k)
            {
                p.fringe.push_back(Pass{dm, dn, p.k - dk, int(p.h_prods.size())});
                p.h_prods.push_back(single_desc(0, dk, dk, 0, 0, 0, false));
            }
        }
        // combine tables for AB / Naive: per C sub-block, the products feeding it in product order
        p.nsub = int(p.d * p.d);
        p.h_term_begin.assign(size_t(p.nsub) + 1, 0);
        for (int s = 0; s < p.nsub; s++)
        {
            p.h_sub_roff.push_back((s / p.d) * p.ms);
            p.h_sub_coff.push_back((s % p.d) * p.ns);
            p.h_term_begin[size_t(s)] = int(p.h_terms.size());
            const auto& prods = products_for_level(level);
            for (size_t q = 0; q < prods.size(); q++)
                for (const SubTerm& t : prods[q].C)
                    if (t.subid == s) p.h_terms.push_back(CombineTerm{int(q), t.coeff});
        }
        p.h_term_begin[size_t(p.nsub)] = int(p.h_terms.size());
        // one-pass form: each (product, sub-block) pair at most once with coefficient +-1, so that visiting
        // the products in order sums every sub-block's terms in its list order
        build_comb_mask(p);
        for (int q = 0; q < p.nprod; q++)
        {
            ProdDesc pd;
            std::memset(&pd, 0, sizeof(pd));
            pd.nA = pd.nB = 1;
            pd.a[0] = SubRef{int64_t(q) * p.ms * p.ks, 0, 0, 1.0};
            pd.b[0] = SubRef{int64_t(q) * p.ks * p.ns, 0, 0, 1.0};
            p.h_dense_prods.push_back(pd);
        }
    }
    p.h_a_m = std::move(a_m);
    p.h_a_k = std::move(a_k);
    p.h_b_k = std::move(b_k);
    p.h_b_n = std::move(b_n);
    p.h_c_m = std::move(c_m);
    p.h_c_n = std::move(c_n);
    p.a_k_step = affine_step(p.h_a_k, p.ks);
    p.b_k_step = affine_step(p.h_b_k, p.ks);
    choose_outer_staged(p, flags);
    // Staged ABC whose stage launches are all single-term (level 1 staged, level 2 all staged) over leaves
    // at most 2048 deep: 64 x 64 tiles, whose four resident workgroups per CU overlap the multi-C epilogues
    // of short leaves better than two of 128 x 128 (measured eff. TFLOP/s: m = n = 16384, k = 1024 level 1
    // 61.6 against 56.8 and level 2 39.7 against 35.6; FP32 4096^3 level 1 forward + adjoints 113.3
    // against 107.5; 8192^3 level 2 level, 74.2 against 74.1; at 4096-deep leaves 128 x 128 stays ahead,
    // 75.2 against 72.1)
    if (!p.tile_given && forced_tile() < 0 && p.core && p.impl == IMPL_ABC && p.outer_staged && p.ks <= 2048 &&
        p.tile == TILE_128)
    {
        bool single = !p.h_inner.empty();
        for (const ProdDesc& q : p.h_inner) single = single && q.nA <= 1 && q.nB <= 1;
        if (single) p.tile = TILE_64;
    }
    // the low-precision leaves stage every sum of a Naive plan (depth = level) in one pass when they can
    if (p.core && p.impl == IMPL_NAIVE && !p.outer_staged) build_msum(p, p.level, p.ms, p.ns, p.ks);
    return Status();
}

// Launch stages for ABC: products of one stage write disjoint C sub-blocks, so a stage is one launch;
// the stages run in order, and the first writer of each C sub-block in that order applies beta.
void make_stages(Plan& p, std::vector<int>& order)
{
    // C target of a term: (staging base, row offset, column offset)
    std::vector<std::pair<int64_t, std::pair<int64_t, int64_t>>> keys;
    std::vector<std::vector<size_t>> targets(size_t(p.nprod));
    for (int q = 0; q < p.nprod; q++)
    {
        const ProdDesc& pd = p.h_prods[size_t(q)];
        for (int t = 0; t < pd.nC; t++)
        {
            const auto key = std::make_pair(pd.c[t].base, std::make_pair(pd.c[t].roff, pd.c[t].coff));
            size_t i = 0;
            while (i < keys.size() && keys[i] != key) i++;
            if (i == keys.size()) keys.push_back(key);
            targets[size_t(q)].push_back(i);
        }
    }
    // First-fit decreasing: products with more C targets first, each into the first stage whose
    // targets it does not share.  For the 1- and 2-level tables this gives 4 stages of {1,2,2,2}
    // products and 16 stages of {1, 2 x6, 4 x9} products (each 2-level stage is the Kronecker product
    // of two 1-level stages), the fewest launches whose workgroup count is an even number of products.
    std::vector<int> by_targets(size_t(p.nprod));
    for (int q = 0; q < p.nprod; q++) by_targets[size_t(q)] = q;
    std::stable_sort(by_targets.begin(), by_targets.end(),
                     [&](int x, int y) { return targets[size_t(x)].size() > targets[size_t(y)].size(); });
    std::vector<std::vector<int>> stages;
    std::vector<std::vector<char>> used;
    for (int q : by_targets)
    {
        size_t st = 0;
        for (; st < stages.size(); st++)
        {
            bool free = true;
            for (size_t i : targets[size_t(q)]) free = free && !used[st][i];
            if (free) break;
        }
        if (st == stages.size())
        {
            stages.emplace_back();
            used.emplace_back(keys.size(), 0);
        }
        stages[st].push_back(q);
        for (size_t i : targets[size_t(q)]) used[st][i] = 1;
    }
    order.clear();
    p.h_stage_begin.assign(1, 0);
    p.h_stage_terms.clear();
    for (const std::vector<int>& s : stages)
    {
        int terms = 1;
        for (int q : s)
        {
            order.push_back(q);
            terms = std::max(terms, max_terms(&p.h_prods[size_t(q)], 1));
        }
        p.h_stage_begin.push_back(int(order.size()));
        p.h_stage_terms.push_back(terms);
    }
    // beta goes with the first writer of each C target in stage order
    p.h_stage_first.assign(size_t(p.nprod), 0);
    std::vector<char> seen(keys.size(), 0);
    for (size_t j = 0; j < order.size(); j++)
    {
        const int q = order[j];
        int mask = 0;
        for (size_t t = 0; t < targets[size_t(q)].size(); t++)
            if (!seen[targets[size_t(q)][t]])
            {
                mask |= 1 << int(t);
                seen[targets[size_t(q)][t]] = 1;
            }
        p.h_stage_first[j] = mask;
    }
}

// Tables of the single persistent launch over every staged product (StageArgs): the products in launch
// order with their first-writer masks for that order (stage_pdesc), per product the counter block of
// each C target (one counter per tile of the grid), how many earlier products update that target (the
// tile waits until they all have), and whether a later one reads it; the tile grid split into one block
// per XCD, as square as the grid allows (an XCD's operand panel reads grow with its block's rows plus
// columns).
//
// Launch order: a tile that has to wait holds its workgroup idle, so the products are ordered greedily
// by when their targets are free in a simple model of the launch (the resident workgroups take tiles in
// queue order at a steady rate; every tile takes one unit of time): next comes the product whose last
// earlier writer of any of its targets finishes first, ties in stage order.
void build_stage_persist(Plan& p, const std::vector<ProdDesc>& staged)
{
    p.stage_persist = false;
    p.h_stage_prods.clear();
    p.h_stage_pdesc.clear();
    p.stage_targets = 0;
    std::memset(&p.h_stage_args, 0, sizeof(p.h_stage_args));
    const int n = int(staged.size());
    // (the persistent stage kernel has the square tiles only)
    if (!p.outer_staged || n < 2 || p.kf_inner.mode != KMODE_AFFINE_VEC || tile_is_rect(p.tile)) return;
    for (int t : p.h_inner_stage_terms)
        if (t != 1) return;
    const int64_t te = tile_extent(p.tile);
    const int64_t tm = (p.ms + te - 1) / te, tn = (p.ns + te - 1) / te;
    // jobs and tile-counter indices stay far below 2^31 (the queue also counts one failed take per
    // workgroup at the end)
    if (tm * tn * n >= (int64_t(1) << 30)) return;
    // targets of every product
    std::vector<std::tuple<int64_t, int64_t, int64_t>> keys;
    std::vector<std::vector<int>> tgt(static_cast<size_t>(n));
    for (int j = 0; j < n; j++)
    {
        const ProdDesc& pd = staged[size_t(j)];
        for (int t = 0; t < pd.nC; t++)
        {
            const auto key = std::make_tuple(pd.c[t].base, pd.c[t].roff, pd.c[t].coff);
            size_t i = 0;
            while (i < keys.size() && keys[i] != key) i++;
            if (i == keys.size()) keys.push_back(key);
            tgt[size_t(j)].push_back(int(i));
        }
    }
    // launch order (see above); 512 resident workgroups (two per CU on 256 CUs)
    std::vector<int> order;
    {
        const double step = double(tm * tn) / 512.0;  // time to take one product's tiles
        std::vector<double> free_at(keys.size(), 0.0);
        std::vector<char> used(size_t(n), 0);
        double now = 0.0;
        for (int k = 0; k < n; k++)
        {
            int best = -1;
            double best_ready = 0.0;
            for (int j = 0; j < n; j++)
            {
                if (used[size_t(j)]) continue;
                double ready = 0.0;
                for (int i : tgt[size_t(j)]) ready = std::max(ready, free_at[size_t(i)]);
                if (best < 0 || ready < best_ready)
                {
                    best = j;
                    best_ready = ready;
                }
            }
            used[size_t(best)] = 1;
            order.push_back(best);
            const double end = std::max(now, best_ready) + step + 1.0;
            for (int i : tgt[size_t(best)]) free_at[size_t(i)] = end;
            now += step;
        }
    }
    // products in launch order, first writers and dependencies for that order
    std::vector<unsigned> writers(keys.size(), 0);
    std::vector<int> last(keys.size(), -1);
    for (int k = 0; k < n; k++)
    {
        const int j = order[size_t(k)];
        ProdDesc pd = staged[size_t(j)];
        StageProd sp;
        std::memset(&sp, 0, sizeof(sp));
        sp.nC = pd.nC;
        pd.first_mask = 0;
        for (int t = 0; t < pd.nC; t++)
        {
            const int i = tgt[size_t(j)][size_t(t)];
            if (writers[size_t(i)] == 0) pd.first_mask |= 1 << t;
            sp.cbase[t] = int64_t(i) * tm * tn;
            sp.need[t] = writers[size_t(i)]++;
            last[size_t(i)] = k;
        }
        p.h_stage_pdesc.push_back(pd);
        p.h_stage_prods.push_back(sp);
    }
    for (int k = 0; k < n; k++)
        for (int i : tgt[size_t(order[size_t(k)])]) p.h_stage_prods[size_t(k)].release |= last[size_t(i)] > k ? 1 : 0;
    StageArgs& a = p.h_stage_args;
    a.nprod = n;
    a.tiles_m = int(tm);
    a.tiles_n = int(tn);
    int64_t best = -1;
    for (int pr = 1; pr <= kStageXcds; pr *= 2)
    {
        const int pc = kStageXcds / pr;
        const int64_t cost = (tm + pr - 1) / pr + (tn + pc - 1) / pc;
        if (best < 0 || cost < best)
        {
            best = cost;
            a.pr = pr;
            a.pc = pc;
        }
    }
    for (int i = 0; i <= a.pr; i++) a.rb[i] = int(tm * i / a.pr);
    for (int i = 0; i <= a.pc; i++) a.cb[i] = int(tn * i / a.pc);
    p.stage_targets = int(keys.size());
    p.stage_persist = true;
}

// Builds the device image of a plan (the byte layout of every table) on the host; bind_image then points
// the plan's device-table members into a copy of it at `base`.
void build_image(Plan& p, PlanImage& img)
{
    const int TK = tile_kdepth(p.tile, p.elem);
    std::vector<int64_t> a_krun = run_strides(p.h_a_k, TK), b_krun = run_strides(p.h_b_k, TK);
    p.kf_core = kfast_check(p, p.h_prods, 0, p.nprod, p.ms, p.ns, p.ks, TK, a_krun, b_krun);
    for (Pass& ps : p.fringe)
        ps.kf = kfast_check(p, p.h_prods, ps.prod_index, 1, ps.ms, ps.ns, ps.ks, TK, a_krun, b_krun);
    std::vector<ProdDesc> fast_prods;
    for (const ProdDesc& pd : p.h_prods) fast_prods.push_back(shifted(p, pd));
    std::vector<int> order;
    make_stages(p, order);
    p.h_stage_order = order;
    std::vector<ProdDesc> staged, fast_staged;
    std::vector<ProdDesc> inner_staged;
    for (size_t j = 0; j < order.size(); j++)
    {
        staged.push_back(p.h_prods[size_t(order[j])]);
        fast_staged.push_back(fast_prods[size_t(order[j])]);
        staged.back().first_mask = fast_staged.back().first_mask = p.h_stage_first[j];
        if (p.outer_staged)
        {
            inner_staged.push_back(p.h_inner[size_t(order[j])]);
            inner_staged.back().first_mask = p.h_stage_first[j];
        }
    }
    p.h_inner_stage_terms.clear();
    for (size_t st = 0; p.outer_staged && st + 1 < p.h_stage_begin.size(); st++)
        p.h_inner_stage_terms.push_back(
            max_terms(inner_staged.data() + p.h_stage_begin[st], p.h_stage_begin[st + 1] - p.h_stage_begin[st]));
    build_stage_persist(p, inner_staged);
    // dense K-fast index vectors of the outer-staged buffers
    std::vector<int64_t> os_sm, os_sk, os_tk, os_tn, os_skrun, os_tkrun;
    for (int64_t i = 0; i < p.os_m; i++) os_sm.push_back(i * p.os_k);
    for (int64_t i = 0; i < p.os_k; i++) os_sk.push_back(i);
    for (int64_t i = 0; i < p.os_k; i++) os_tk.push_back(i);
    for (int64_t i = 0; i < p.os_n; i++) os_tn.push_back(i * p.os_k);
    if (p.outer_staged)
    {
        os_skrun = run_strides(os_sk, TK);
        os_tkrun = run_strides(os_tk, TK);
    }
    if (p.core)
    {
        // dense S_p (ms x ks, column-major) and T_p (ks x ns): K strides ms and 1, terms at base offsets
        const int64_t limit = (int64_t(1) << 31) / p.elem;
        p.kf_dense.on = (p.ms - 1) + (TK - 1) * p.ms < limit && (p.ns - 1) * p.ks + (TK - 1) < limit;
        p.kf_dense.mode = !p.kf_dense.on ? KMODE_GENERAL : (p.ks % TK == 0 ? KMODE_AFFINE_VEC : KMODE_AFFINE);
        p.kf_dense.sa = p.ms;
        p.kf_dense.sb = 1;
        // S K-fast ([q][i][k]): both operands K-fast, the direct-to-LDS path when K-tiles are whole
        p.kf_dense_k.on = (p.ms - 1) * p.ks + TK < limit && (p.ns - 1) * p.ks + TK < limit;
        p.kf_dense_k.mode = !p.kf_dense_k.on ? KMODE_GENERAL : (p.ks % TK == 0 ? KMODE_AFFINE_VEC : KMODE_AFFINE);
        p.kf_dense_k.sa = 1;
        p.kf_dense_k.sb = 1;
    }
    std::vector<int64_t> s_m, s_k, t_k, t_n, s_krun, t_krun, sk_m;
    if (p.core)
    {
        for (int64_t i = 0; i < p.ms; i++) s_m.push_back(i);
        for (int64_t i = 0; i < p.ms; i++) sk_m.push_back(i * p.ks);
        for (int64_t i = 0; i < p.ks; i++) s_k.push_back(i * p.ms);
        for (int64_t i = 0; i < p.ks; i++) t_k.push_back(i);
        for (int64_t i = 0; i < p.ns; i++) t_n.push_back(i * p.ks);
        s_krun = run_strides(s_k, TK);
        t_krun = run_strides(t_k, TK);
    }

    struct Item
    {
        const void* src;
        size_t bytes;
        size_t off;
    };
    std::vector<Item> items;
    size_t total = 0;
    auto add = [&](const void* src, size_t bytes) {
        items.push_back(Item{src, bytes, total});
        total += align256(bytes > 0 ? bytes : 8);
        return items.size() - 1;
    };
    auto i_a_m = add(p.h_a_m.data(), p.h_a_m.size() * 8), i_a_k = add(p.h_a_k.data(), p.h_a_k.size() * 8),
         i_a_kr = add(a_krun.data(), a_krun.size() * 8), i_b_k = add(p.h_b_k.data(), p.h_b_k.size() * 8),
         i_b_kr = add(b_krun.data(), b_krun.size() * 8), i_b_n = add(p.h_b_n.data(), p.h_b_n.size() * 8),
         i_c_m = add(p.h_c_m.data(), p.h_c_m.size() * 8), i_c_n = add(p.h_c_n.data(), p.h_c_n.size() * 8),
         i_ba = add(p.h_batch_a.data(), p.h_batch_a.size() * 8),
         i_bb = add(p.h_batch_b.data(), p.h_batch_b.size() * 8),
         i_bc = add(p.h_batch_c.data(), p.h_batch_c.size() * 8),
         i_pr = add(p.h_prods.data(), p.h_prods.size() * sizeof(ProdDesc)),
         i_prf = add(fast_prods.data(), fast_prods.size() * sizeof(ProdDesc)),
         i_pst = add(staged.data(), staged.size() * sizeof(ProdDesc)),
         i_pfst = add(fast_staged.data(), fast_staged.size() * sizeof(ProdDesc)),
         i_sr = add(p.h_sub_roff.data(), p.h_sub_roff.size() * 8),
         i_sc = add(p.h_sub_coff.data(), p.h_sub_coff.size() * 8),
         i_tb = add(p.h_term_begin.data(), p.h_term_begin.size() * sizeof(int)),
         i_te = add(p.h_terms.data(), p.h_terms.size() * sizeof(CombineTerm)),
         i_sm = add(s_m.data(), s_m.size() * 8), i_sk = add(s_k.data(), s_k.size() * 8),
         i_skm = add(sk_m.data(), sk_m.size() * 8),
         i_skr = add(s_krun.data(), s_krun.size() * 8), i_tk = add(t_k.data(), t_k.size() * 8),
         i_tkr = add(t_krun.data(), t_krun.size() * 8), i_tn = add(t_n.data(), t_n.size() * 8),
         i_dp = add(p.h_dense_prods.data(), p.h_dense_prods.size() * sizeof(ProdDesc)),
         i_zero = add(nullptr, 256),
         i_os = add(p.h_outer_sums.data(), p.h_outer_sums.size() * sizeof(ProdDesc)),
         i_mca = add(p.h_msum_coef_a.data(), p.h_msum_coef_a.size() * 8),
         i_mcb = add(p.h_msum_coef_b.data(), p.h_msum_coef_b.size() * 8),
         i_msa = add(p.h_msum_shift_a.data(), p.h_msum_shift_a.size() * 8),
         i_mc1a = add(p.h_msum_coef1_a.data(), p.h_msum_coef1_a.size() * 8),
         i_mc1b = add(p.h_msum_coef1_b.data(), p.h_msum_coef1_b.size() * 8),
         i_msb = add(p.h_msum_shift_b.data(), p.h_msum_shift_b.size() * 8),
         i_osf = add(p.h_outer_shifted.data(), p.h_outer_shifted.size() * sizeof(ProdDesc)),
         i_ins = add(inner_staged.data(), inner_staged.size() * sizeof(ProdDesc)),
         i_osm = add(os_sm.data(), os_sm.size() * 8), i_osk = add(os_sk.data(), os_sk.size() * 8),
         i_oskr = add(os_skrun.data(), os_skrun.size() * 8), i_otk = add(os_tk.data(), os_tk.size() * 8),
         i_otkr = add(os_tkrun.data(), os_tkrun.size() * 8), i_otn = add(os_tn.data(), os_tn.size() * 8),
         i_spp = add(p.h_stage_prods.data(), p.h_stage_prods.size() * sizeof(StageProd)),
         i_spd = add(p.h_stage_pdesc.data(), p.h_stage_pdesc.size() * sizeof(ProdDesc));

    img.bytes.assign(total, 0);
    for (const Item& it : items)
        if (it.bytes && it.src) std::memcpy(img.bytes.data() + it.off, it.src, it.bytes);
    const size_t idx[] = {i_a_m, i_a_k, i_a_kr, i_b_k, i_b_kr, i_b_n, i_c_m, i_c_n, i_ba, i_bb, i_bc, i_pr, i_prf,
                          i_pst, i_pfst, i_sr, i_sc, i_tb, i_te, i_sm, i_sk, i_skm, i_skr, i_tk, i_tkr, i_tn, i_dp,
                          i_zero, i_os, i_mca, i_mcb, i_msa, i_mc1a, i_mc1b, i_msb, i_osf, i_ins, i_osm, i_osk,
                          i_oskr, i_otk, i_otkr, i_otn, i_spp, i_spd};
    static_assert(sizeof(idx) / sizeof(idx[0]) == PlanImage::kParts, "one offset per table");
    for (int i = 0; i < PlanImage::kParts; i++) img.off[i] = items[idx[i]].off;
}

// Points the plan's device tables at `d`, which holds a copy of the image build_image made.
void bind_image(Plan& p, const PlanImage& img, void* d)
{
    // the parts in the order build_image records their offsets
    enum
    {
        i_a_m, i_a_k, i_a_kr, i_b_k, i_b_kr, i_b_n, i_c_m, i_c_n, i_ba, i_bb, i_bc, i_pr, i_prf, i_pst, i_pfst,
        i_sr, i_sc, i_tb, i_te, i_sm, i_sk, i_skm, i_skr, i_tk, i_tkr, i_tn, i_dp, i_zero, i_os, i_mca, i_mcb,
        i_msa, i_mc1a, i_mc1b, i_msb, i_osf, i_ins, i_osm, i_osk, i_oskr, i_otk, i_otkr, i_otn, i_spp, i_spd, n_parts
    };
    static_assert(n_parts == PlanImage::kParts, "one offset per table");
    auto at = [&](size_t i) { return reinterpret_cast<unsigned char*>(d) + img.off[i]; };
    p.a_m = reinterpret_cast<const int64_t*>(at(i_a_m));
    p.a_k = reinterpret_cast<const int64_t*>(at(i_a_k));
    p.a_krun = reinterpret_cast<const int64_t*>(at(i_a_kr));
    p.b_k = reinterpret_cast<const int64_t*>(at(i_b_k));
    p.b_krun = reinterpret_cast<const int64_t*>(at(i_b_kr));
    p.b_n = reinterpret_cast<const int64_t*>(at(i_b_n));
    p.c_m = reinterpret_cast<const int64_t*>(at(i_c_m));
    p.c_n = reinterpret_cast<const int64_t*>(at(i_c_n));
    p.batch_a = reinterpret_cast<const int64_t*>(at(i_ba));
    p.batch_b = reinterpret_cast<const int64_t*>(at(i_bb));
    p.batch_c = reinterpret_cast<const int64_t*>(at(i_bc));
    p.prods = reinterpret_cast<const ProdDesc*>(at(i_pr));
    p.prods_fast = reinterpret_cast<const ProdDesc*>(at(i_prf));
    p.prods_staged = reinterpret_cast<const ProdDesc*>(at(i_pst));
    p.prods_fast_staged = reinterpret_cast<const ProdDesc*>(at(i_pfst));
    p.sub_roff = reinterpret_cast<const int64_t*>(at(i_sr));
    p.sub_coff = reinterpret_cast<const int64_t*>(at(i_sc));
    p.term_begin = reinterpret_cast<const int*>(at(i_tb));
    p.terms = reinterpret_cast<const CombineTerm*>(at(i_te));
    p.s_m = reinterpret_cast<const int64_t*>(at(i_sm));
    p.sk_m = reinterpret_cast<const int64_t*>(at(i_skm));
    p.s_k = reinterpret_cast<const int64_t*>(at(i_sk));
    p.s_krun = reinterpret_cast<const int64_t*>(at(i_skr));
    p.t_k = reinterpret_cast<const int64_t*>(at(i_tk));
    p.t_krun = reinterpret_cast<const int64_t*>(at(i_tkr));
    p.t_n = reinterpret_cast<const int64_t*>(at(i_tn));
    p.dense_prods = reinterpret_cast<const ProdDesc*>(at(i_dp));
    p.zero = at(i_zero);
    p.outer_sums = reinterpret_cast<const ProdDesc*>(at(i_os));
    p.msum_coef_a = reinterpret_cast<const double*>(at(i_mca));
    p.msum_coef_b = reinterpret_cast<const double*>(at(i_mcb));
    p.msum_shift_a = reinterpret_cast<const int64_t*>(at(i_msa));
    p.msum_coef1_a = p.h_msum_coef1_a.empty() ? nullptr : reinterpret_cast<const double*>(at(i_mc1a));
    p.msum_coef1_b = p.h_msum_coef1_b.empty() ? nullptr : reinterpret_cast<const double*>(at(i_mc1b));
    p.msum_shift_b = reinterpret_cast<const int64_t*>(at(i_msb));
    p.outer_shifted = reinterpret_cast<const ProdDesc*>(at(i_osf));
    p.inner_staged = reinterpret_cast<const ProdDesc*>(at(i_ins));
    p.os_sm = reinterpret_cast<const int64_t*>(at(i_osm));
    p.os_sk = reinterpret_cast<const int64_t*>(at(i_osk));
    p.os_skrun = reinterpret_cast<const int64_t*>(at(i_oskr));
    p.os_tk = reinterpret_cast<const int64_t*>(at(i_otk));
    p.os_tkrun = reinterpret_cast<const int64_t*>(at(i_otkr));
    p.os_tn = reinterpret_cast<const int64_t*>(at(i_otn));
    p.h_stage_args.prods = reinterpret_cast<const StageProd*>(at(i_spp));
    p.stage_pdesc = reinterpret_cast<const ProdDesc*>(at(i_spd));
}

namespace {

Status upload(Plan& p)
{
    PlanImage img;
    build_image(p, img);
    void* d = nullptr;
    hipError_t e = hipMalloc(&d, img.bytes.size());
    if (e != hipSuccess) return Status::error(MISTRA_ERR_NOMEM, std::string("plan allocation: ") + hipGetErrorString(e));
    e = hipMemcpy(d, img.bytes.data(), img.bytes.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess)
    {
        (void)hipFree(d);
        return Status::error(MISTRA_ERR_HIP, std::string("plan upload: ") + hipGetErrorString(e));
    }
    p.dmem = d;
    bind_image(p, img, d);
    return Status();
}

}  // namespace

std::string plan_key(const tblis::tblis_tensor* A, const label_type* idx_A, const tblis::tblis_tensor* B,
                     const label_type* idx_B, const tblis::tblis_tensor* C, const label_type* idx_C, int level,
                     int impl, int tile, unsigned flags, int dev)
{
    std::ostringstream ks;
    ks << dev << '|' << int(A->type) << '|' << level << '|' << impl << '|' << tile << '|'
       << (flags & (MISTRA_FLAG_REF_FRINGE | MISTRA_FLAG_FUSED_SUMS | MISTRA_FLAG_OUTER_STAGED |
                    MISTRA_FLAG_ALL_STAGED))
       << '|';
    std::string key = ks.str();
    append_key(key, A, idx_A);
    append_key(key, B, idx_B);
    append_key(key, C, idx_C);
    return key;
}

Status get_plan(const tblis::tblis_tensor* A, const label_type* idx_A, const tblis::tblis_tensor* B,
                const label_type* idx_B, const tblis::tblis_tensor* C, const label_type* idx_C, int level,
                int impl, int tile, unsigned flags, std::shared_ptr<Plan>& out, bool capturing)
{
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    const std::string key = plan_key(A, idx_A, B, idx_B, C, idx_C, level, impl, tile, flags, dev);
    {
        std::lock_guard<std::mutex> lk(g_plan_mutex);
        auto f = g_plan_index.find(key);
        if (f != g_plan_index.end())
        {
            out = f->second->second;
            g_plans.splice(g_plans.begin(), g_plans, f->second);
            if (capturing) g_pinned.emplace(key, out);
            return Status();
        }
        auto pin = g_pinned.find(key);
        if (pin != g_pinned.end())
        {
            out = pin->second;
            return Status();
        }
    }
    // building a plan allocates and uploads its tables synchronously, which a stream capture forbids
    if (capturing)
        return Status::error(MISTRA_ERR_UNSUPPORTED,
                             "no plan for this contraction yet: make one call with the same shapes, strides, level, "
                             "impl and flags outside the graph capture first");
    auto p = std::make_shared<Plan>();
    Status st = build_plan_host(A, idx_A, B, idx_B, C, idx_C, level, impl, tile, flags, *p);
    if (!st.ok()) return st;
    if (!p->general)
    {
        st = upload(*p);
        if (!st.ok()) return st;
    }
    std::lock_guard<std::mutex> lk(g_plan_mutex);
    auto f = g_plan_index.find(key);
    if (f != g_plan_index.end())  // built concurrently by another thread: keep that one
    {
        out = f->second->second;
        return Status();
    }
    g_plans.emplace_front(key, p);
    g_plan_index[key] = g_plans.begin();
    while (g_plans.size() > kMaxPlans)
    {
        g_plan_index.erase(g_plans.back().first);
        g_plans.pop_back();
    }
    out = p;
    return Status();
}

void release_all()
{
    // pinned plans (captured by a caller's graph) are kept: their graphs may still be replayed
    std::lock_guard<std::mutex> lk(g_plan_mutex);
    g_plan_index.clear();
    g_plans.clear();
}

hipError_t stream_alloc(void** out, size_t bytes, hipStream_t stream)
{
    static std::mutex m;
    static std::vector<bool> kept;  // per device: release threshold raised
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    {
        std::lock_guard<std::mutex> lk(m);
        if (size_t(dev) >= kept.size()) kept.resize(size_t(dev) + 1, false);
        if (!kept[size_t(dev)])
        {
            hipMemPool_t pool = nullptr;
            uint64_t keep = UINT64_MAX;
            if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess)
                (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
            (void)hipGetLastError();  // a pool without the attribute keeps the runtime's default
            kept[size_t(dev)] = true;
        }
    }
    return hipMallocAsync(out, bytes, stream);
}

void trim_stream_pool()
{
    int dev = 0;
    hipMemPool_t pool = nullptr;
    if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess)
    {
        (void)hipDeviceSynchronize();
        (void)hipMemPoolTrimTo(pool, 0);
    }
    (void)hipGetLastError();
}

namespace {

GemmArgs base_args(const Plan& p, const Call& c)
{
    GemmArgs g;
    std::memset(&g, 0, sizeof(g));
    g.A = c.A;
    g.B = c.B;
    g.C = c.C;
    g.a_m = p.a_m;
    g.a_k = p.a_k;
    g.a_krun = p.a_krun;
    g.b_k = p.b_k;
    g.b_krun = p.b_krun;
    g.b_n = p.b_n;
    g.c_m = p.c_m;
    g.c_n = p.c_n;
    g.alpha = c.alpha;
    g.beta = c.beta;
    g.zero = p.zero;
    g.nbatch = p.nbatch;
    g.batch_a = p.nbatch > 1 ? p.batch_a : nullptr;
    g.batch_b = p.nbatch > 1 ? p.batch_b : nullptr;
    g.batch_c = p.nbatch > 1 ? p.batch_c : nullptr;
    if (p.nbatch == 1)
    {
        g.A = static_cast<const char*>(c.A) + p.h_batch_a[0] * p.elem;
        g.B = static_cast<const char*>(c.B) + p.h_batch_b[0] * p.elem;
        g.C = static_cast<char*>(c.C) + p.h_batch_c[0] * p.elem;
    }
    return g;
}

double leaf_flops(const Plan& p, int nprod, int64_t ms, int64_t ns, int64_t ks)
{
    return 2.0 * double(nprod) * double(ms) * double(ns) * double(ks) * double(p.nbatch);
}

}  // namespace

Status plan_abc_jobs(const Plan& p, const Call& c, std::vector<LaunchJob>& out)
{
    if (p.general) return Status::error(MISTRA_ERR_UNSUPPORTED, "general index classes run outside a batch");
    if (p.outer_staged) return Status::error(MISTRA_ERR_UNSUPPORTED, "outer-staged plans run outside a batch");
    auto push = [&](GemmArgs g, int tile, int terms, int kmode, int round, double flops) {
        if (tile_is_rect(tile)) tile = TILE_128;  // grouped launches have the square tiles only
        tile = effective_tile(tile, kmode, terms, p.a_mfast, p.b_nfast);
        const int te = tile_extent(tile);
        g.tiles_m = int((g.ms + te - 1) / te);
        g.tiles_n = int((g.ns + te - 1) / te);
        if (g.tiles_m == 0 || g.tiles_n == 0 || g.nbatch == 0 || g.nprod == 0 || g.ks <= 0) return;
        out.push_back(LaunchJob{g, p.elem, tile, terms, round, kmode, p.a_mfast, p.b_nfast, flops});
    };
    int round = 0;
    if (p.nprod > 0)
    {
        GemmArgs g = base_args(p, c);
        g.ms = p.ms;
        g.ns = p.ns;
        g.ks = p.ks;
        g.ka_stride = p.kf_core.sa;
        g.kb_stride = p.kf_core.sb;
        for (size_t st = 0; st + 1 < p.h_stage_begin.size(); st++, round++)
        {
            const int b = p.h_stage_begin[st], n = p.h_stage_begin[st + 1] - b;
            g.prods = (p.kf_core.on ? p.prods_fast_staged : p.prods_staged) + b;
            g.nprod = n;
            push(g, p.tile, p.h_stage_terms[st], p.kf_core.mode, round, leaf_flops(p, n, p.ms, p.ns, p.ks));
        }
    }
    for (const Pass& ps : p.fringe)
    {
        GemmArgs f = base_args(p, c);
        f.ms = ps.ms;
        f.ns = ps.ns;
        f.ks = ps.ks;
        f.prods = (ps.kf.on ? p.prods_fast : p.prods) + ps.prod_index;
        f.nprod = 1;
        f.ka_stride = ps.kf.sa;
        f.kb_stride = ps.kf.sb;
        push(f, TILE_64, 1, ps.kf.mode, round++, leaf_flops(p, 1, ps.ms, ps.ns, ps.ks));
    }
    return Status();
}

namespace {

// A second stream of the current device, for an independent pass beside the caller's stream: the sums
// of A and of B read and write disjoint buffers, and each alone leaves HBM bandwidth unused.  One per
// thread and device, created on first use with the default priority and kept for the thread's life.
struct SideFork
{
    hipStream_t side = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
};

// The side stream waits for everything issued to `main` so far.  false (run everything on `main`) when
// the stream is being captured into a graph or a call fails.
bool fork_begin(hipStream_t main, SideFork& f)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    int dev = 0;
    if (hipStreamIsCapturing(main, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone ||
        hipGetDevice(&dev) != hipSuccess)
    {
        (void)hipGetLastError();
        return false;
    }
    thread_local std::map<int, SideFork> forks;
    auto it = forks.find(dev);
    if (it == forks.end())
    {
        SideFork n;
        if (hipStreamCreateWithFlags(&n.side, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&n.ev[0], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&n.ev[1], hipEventDisableTiming) != hipSuccess)
        {
            (void)hipGetLastError();
            if (n.ev[0]) (void)hipEventDestroy(n.ev[0]);
            if (n.side) (void)hipStreamDestroy(n.side);
            return false;
        }
        it = forks.emplace(dev, n).first;
    }
    f = it->second;
    if (hipEventRecord(f.ev[0], main) != hipSuccess || hipStreamWaitEvent(f.side, f.ev[0], 0) != hipSuccess)
    {
        (void)hipGetLastError();
        return false;
    }
    return true;
}

// `main` waits for everything issued to the side stream so far.
hipError_t fork_end(hipStream_t main, const SideFork& f)
{
    hipError_t e = hipEventRecord(f.ev[1], f.side);
    if (e == hipSuccess) e = hipStreamWaitEvent(main, f.ev[1], 0);
    return e;
}

// Workspace of a call: from its arena when it has one (hipErrorOutOfMemory when the block is too small),
// else from the stream-ordered pool.
hipError_t ws_alloc(const Call& c, void** out, size_t bytes, hipStream_t s)
{
    if (!c.ws) return stream_alloc(out, bytes, s);
    const size_t off = align256(c.ws->used);
    if (off > c.ws->size || bytes > c.ws->size - off) return hipErrorOutOfMemory;
    *out = c.ws->base + off;
    c.ws->used = off + bytes;
    return hipSuccess;
}

hipError_t ws_free(const Call& c, void* p, hipStream_t s)
{
    return c.ws ? hipSuccess : hipFreeAsync(p, s);
}

}  // namespace

Status stage_wait_status()
{
    unsigned n = 0;
    const hipError_t e = stage_timeouts_take(&n);
    if (e != hipSuccess) return Status::error(MISTRA_ERR_HIP, std::string("stage launch wait count: ") + hipGetErrorString(e));
    if (n)
        return Status::error(MISTRA_ERR_HIP, "persistent stage launch: " + std::to_string(n) +
                                                 " tile waits gave up (results are not valid)");
    return Status();
}

namespace {

// Polls a stage-launch wait makes before it gives up (each a coherent load and a short sleep: seconds)
constexpr long long kStageSpin = 1ll << 22;

// Register tile of the atomic one-launch form: the plan's when the caller or the environment named one,
// else 64x128 (c2, 8192^3: L2 77.3 eff. TFLOP/s against 76.2 with 64x64, 75.6 with 128x128, 75.9 with
// 128x64; L1 75.1 against 75.9 with 128x128, whose 4096-deep leaves keep the square tile ahead).
int atomic_tile(const Plan& p)
{
    if (p.tile_given || forced_tile() >= 0) return p.tile;
    return p.ks <= 2048 ? TILE_64x128 : TILE_128;
}

// Outer-staged ABC (Plan::outer_staged): the outer sums into S and T, then the products over the inner
// sums of S and T (2 terms at level 2 with one staged level, else 1) with the multi-C epilogue on C: one
// launch per stage, or with MISTRA_FLAG_STAGE_PERSIST (when every product has one term per operand,
// Plan::stage_persist) all of them in one persistent launch.
hipError_t run_outer_staged(const Plan& p, const Call& c, hipStream_t stream)
{
    const size_t nouter = p.h_outer_sums.size();
    const size_t sbytes = nouter * size_t(p.os_m) * size_t(p.os_k) * size_t(p.elem);
    const size_t tbytes = nouter * size_t(p.os_k) * size_t(p.os_n) * size_t(p.elem);
    void *S = nullptr, *T = nullptr;
    // an allocation failure returns before anything is launched (the caller then takes the fused form)
    hipError_t e = ws_alloc(c, &S, sbytes, stream);
    if (e != hipSuccess) return e;
    e = ws_alloc(c, &T, tbytes, stream);
    if (e != hipSuccess)
    {
        (void)ws_free(c, S, stream);
        return e;
    }
    // sums of one operand: one pass over its source sub-blocks when they are translates and K (the rows
    // of the staged buffers) has unit stride, else one gather pass per sum
    const uintptr_t align = uintptr_t(2 * p.elem) - 1;
    auto sums = [&](int which, hipStream_t st) {
        const bool isA = which == 0;
        const void* X = isA ? c.A : c.B;
        const bool rows_fast = isA ? !p.a_mfast : !p.b_nfast;
        const bool pair = (isA ? p.outer_pair_a : p.outer_pair_b) && (reinterpret_cast<uintptr_t>(X) & align) == 0;
        if ((isA ? p.msum_a : p.msum_b) && rows_fast)
        {
            MultiSumArgs ma;
            std::memset(&ma, 0, sizeof(ma));
            ma.X = X;
            ma.zoff = isA ? p.h_batch_a[0] : p.h_batch_b[0];
            ma.rs = isA ? p.a_k : p.b_k;
            ma.cs = isA ? p.a_m : p.b_n;
            ma.rows = p.os_k;
            ma.cols = isA ? p.os_m : p.os_n;
            ma.nsrc = p.msum_nsrc;
            ma.nout = int(nouter);
            ma.coef = isA ? p.msum_coef_a : p.msum_coef_b;
            ma.coef1 = isA ? p.msum_coef1_a : p.msum_coef1_b;
            set_msum_cv(p, isA, ma);
            ma.shift = isA ? p.msum_shift_a : p.msum_shift_b;
            ma.out = isA ? S : T;
            ma.pair = pair;
            return launch_operand_multisums(p.elem, ma, st);
        }
        SumArgs sa;
        std::memset(&sa, 0, sizeof(sa));
        sa.X = X;
        sa.zoff = isA ? p.h_batch_a[0] : p.h_batch_b[0];
        sa.rs = isA ? p.a_k : p.b_k;
        sa.cs = isA ? p.a_m : p.b_n;
        sa.rows = p.os_k;
        sa.cols = isA ? p.os_m : p.os_n;
        sa.nprod = int(nouter);
        sa.which = which;
        sa.rows_fast = rows_fast;
        sa.out = isA ? S : T;
        // translate form (shifted bases) only in the kernel that reads along rows
        sa.shifted = (isA ? p.outer_shift_a : p.outer_shift_b) && rows_fast;
        sa.pair = pair;
        sa.prods = sa.shifted ? p.outer_shifted : p.outer_sums;
        return launch_operand_sums(p.elem, sa, st);
    };
    // B's sums on the side stream, A's on the caller's, joined before the stage launches
    SideFork fk;
    const bool forked = !c.ws && fork_begin(stream, fk);
    e = sums(0, stream);
    const hipError_t eb = sums(1, forked ? fk.side : stream);
    if (e == hipSuccess) e = eb;
    if (forked)
    {
        const hipError_t ej = fork_end(stream, fk);
        if (e == hipSuccess) e = ej;
    }
    GemmArgs g;
    std::memset(&g, 0, sizeof(g));
    g.A = S;
    g.B = T;
    g.C = static_cast<char*>(c.C) + p.h_batch_c[0] * p.elem;
    g.a_m = p.os_sm;
    g.a_k = p.os_sk;
    g.a_krun = p.os_skrun;
    g.b_k = p.os_tk;
    g.b_krun = p.os_tkrun;
    g.b_n = p.os_tn;
    g.c_m = p.c_m;
    g.c_n = p.c_n;
    g.ms = p.ms;
    g.ns = p.ns;
    g.ks = p.ks;
    g.alpha = c.alpha;
    g.beta = c.beta;
    g.nbatch = 1;
    g.zero = p.zero;
    g.ka_stride = p.kf_inner.sa;
    g.kb_stride = p.kf_inner.sb;
    // every stage product in one persistent launch; its counters come from the workspace (a failed
    // allocation leaves the per-stage launches below, which need none)
    bool stages_done = false;
    if (e == hipSuccess && p.stage_persist && (c.flags & MISTRA_FLAG_STAGE_PERSIST))
    {
        const StageArgs& t = p.h_stage_args;
        const int64_t ncnt = stage_counters(t.nprod, p.stage_targets, t.tiles_m, t.tiles_n);
        void* cnt = nullptr;
        if (ws_alloc(c, &cnt, size_t(ncnt) * sizeof(unsigned), stream) == hipSuccess)
        {
            StageArgs a = t;
            a.counters = static_cast<unsigned*>(cnt);
            const long long spin = g_dbg_batch_spin.load();
            a.spin_limit = spin > 0 ? spin : kStageSpin;
            a.dep_bias = g_dbg_batch_bias.load();
            e = stage_timeouts_word(&a.timeouts);
            if (e == hipSuccess) e = launch_zero_u32(a.counters, ncnt, stream);
            int dev = 0, cus = 256;
            if (e == hipSuccess) e = hipGetDevice(&dev);
            if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
            const int64_t jobs = int64_t(t.tiles_m) * t.tiles_n * t.nprod;
            const int grid = int(std::max<int64_t>(1, std::min<int64_t>(int64_t(stage_persistent_occupancy(p.elem, p.tile)) * cus, jobs)));
            GemmArgs gs = g;
            gs.prods = p.stage_pdesc;
            gs.nprod = 1;
            gs.tiles_m = t.tiles_m;
            gs.tiles_n = t.tiles_n;
            gs.c_coherent = 1;
            if (e == hipSuccess)
            {
                timing_begin(stream);
                e = launch_stage_persistent(p.elem, p.tile, a, gs, grid, stream);
                timing_end(stream, leaf_flops(p, t.nprod, p.ms, p.ns, p.ks));
            }
            const hipError_t f = ws_free(c, cnt, stream);
            if (e == hipSuccess) e = f;
            stages_done = true;
        }
        else
            (void)hipGetLastError();
    }
    // MISTRA_FLAG_ATOMIC_C: C's targets scaled by beta once, then every product in one launch adding
    // alpha*coeff*M_p into its targets with FP64 atomic adds (single-term products only)
    bool single_term = p.elem == 8 && !p.h_inner_stage_terms.empty();
    for (int t : p.h_inner_stage_terms) single_term = single_term && t == 1;
    if (!stages_done && e == hipSuccess && single_term && (c.flags & MISTRA_FLAG_ATOMIC_C))
    {
        const int nprod = p.h_stage_begin.back();
        if (c.beta != 1.0)
            e = launch_scale_first_targets(g.C, p.c_m, p.c_n, p.ms, p.ns, p.inner_staged, nprod, c.beta, stream);
        g.prods = p.inner_staged;
        g.nprod = nprod;
        timing_begin(stream);
        if (e == hipSuccess)
            e = launch_leaf_gemm(p.elem, EPI_ATOMIC_C, atomic_tile(p), 1, p.kf_inner.mode, g, false, false, stream);
        timing_end(stream, leaf_flops(p, g.nprod, p.ms, p.ns, p.ks));
        stages_done = true;
    }
    for (size_t st = 0; !stages_done && e == hipSuccess && st + 1 < p.h_stage_begin.size(); st++)
    {
        const int b = p.h_stage_begin[st], n = p.h_stage_begin[st + 1] - b;
        g.prods = p.inner_staged + b;
        g.nprod = n;
        timing_begin(stream);
        e = launch_leaf_gemm(p.elem, EPI_MULTI_C, p.tile, p.h_inner_stage_terms[st], p.kf_inner.mode, g, false,
                             false, stream);
        timing_end(stream, leaf_flops(p, n, p.ms, p.ns, p.ks));
    }
    const hipError_t f1 = ws_free(c, S, stream), f2 = ws_free(c, T, stream);
    if (e != hipSuccess) return e;
    return f1 != hipSuccess ? f1 : f2;
}

}  // namespace

Status run_plan(const Plan& p, const Call& c, hipStream_t stream)
{
    if (p.general) return Status::error(MISTRA_ERR_UNSUPPORTED, "contractions with A-only, B-only or C-only indices are not supported yet");
    if (p.nprod == 0) return Status();
    GemmArgs g = base_args(p, c);

    // staged form; when its workspace cannot be allocated (nothing has run yet), the fused form below,
    // which needs none
    bool staged_done = false;
    if (p.outer_staged)
    {
        const hipError_t e = run_outer_staged(p, c, stream);
        if (e == hipErrorOutOfMemory)
            (void)hipGetLastError();
        else
        {
            HIP_TRY(e);
            staged_done = true;
        }
    }
    if (staged_done)
    {
    }
    else if (!p.core || p.impl == IMPL_ABC || p.impl == IMPL_BLAS || p.impl == IMPL_REF)
    {
        g.ms = p.ms;
        g.ns = p.ns;
        g.ks = p.ks;
        g.ka_stride = p.kf_core.sa;
        g.kb_stride = p.kf_core.sb;
        // one launch per stage (products with disjoint C targets); stream order serializes the stages
        for (size_t st = 0; st + 1 < p.h_stage_begin.size(); st++)
        {
            const int b = p.h_stage_begin[st], n = p.h_stage_begin[st + 1] - b;
            g.prods = (p.kf_core.on ? p.prods_fast_staged : p.prods_staged) + b;
            g.nprod = n;
            timing_begin(stream);
            HIP_TRY(launch_leaf_gemm(p.elem, EPI_MULTI_C, p.tile, p.h_stage_terms[st], p.kf_core.mode, g, p.a_mfast,
                                     p.b_nfast, stream));
            timing_end(stream, leaf_flops(p, n, p.ms, p.ns, p.ks));
        }
    }
    else if (p.impl == IMPL_AB)
    {
        const size_t mbytes = size_t(p.nbatch) * size_t(p.nprod) * size_t(p.ms) * size_t(p.ns) * size_t(p.elem);
        void* M = nullptr;
        {
            const hipError_t ea = ws_alloc(c, &M, mbytes, stream);
            if (ea != hipSuccess)
                return Status::error(MISTRA_ERR_NOMEM, std::string("AB workspace: ") + hipGetErrorString(ea));
        }
        g.ms = p.ms;
        g.ns = p.ns;
        g.ks = p.ks;
        g.prods = p.kf_core.on ? p.prods_fast : p.prods;
        g.nprod = p.nprod;
        g.M = M;
        g.ka_stride = p.kf_core.sa;
        g.kb_stride = p.kf_core.sb;
        timing_begin(stream);
        hipError_t e = launch_leaf_gemm(p.elem, EPI_STORE, p.tile, max_terms(p.h_prods.data(), p.nprod),
                                        p.kf_core.mode, g, p.a_mfast, p.b_nfast, stream);
        timing_end(stream, leaf_flops(p, p.nprod, p.ms, p.ns, p.ks));
        if (e == hipSuccess)
        {
            CombineArgs ca;
            std::memset(&ca, 0, sizeof(ca));
            ca.C = g.C;
            ca.M = M;
            ca.c_m = p.c_m;
            ca.c_n = p.c_n;
            ca.ms = p.ms;
            ca.ns = p.ns;
            ca.nsub = p.nsub;
            ca.sub_roff = p.sub_roff;
            ca.sub_coff = p.sub_coff;
            ca.term_begin = p.term_begin;
            ca.terms = p.terms;
            set_comb_mask(p, ca);
            ca.nprod = p.nprod;
            ca.alpha = c.alpha;
            ca.beta = c.beta;
            ca.nbatch = p.nbatch;
            ca.batch_c = g.batch_c;
            e = launch_combine(p.elem, ca, stream);
        }
        hipError_t f = ws_free(c, M, stream);
        HIP_TRY(e);
        HIP_TRY(f);
    }
    else  // IMPL_NAIVE
    {
        const size_t sbytes = size_t(p.nprod) * size_t(p.ms) * size_t(p.ks) * size_t(p.elem);
        const size_t tbytes = size_t(p.nprod) * size_t(p.ks) * size_t(p.ns) * size_t(p.elem);
        const size_t mbytes = size_t(p.nprod) * size_t(p.ms) * size_t(p.ns) * size_t(p.elem);
        void *S = nullptr, *T = nullptr, *M = nullptr;
        // on a failed allocation the ones already made go back to the pool before the error returns
        hipError_t e = ws_alloc(c, &S, sbytes, stream);
        if (e == hipSuccess) e = ws_alloc(c, &T, tbytes, stream);
        if (e == hipSuccess) e = ws_alloc(c, &M, mbytes, stream);
        if (e != hipSuccess)
        {
            for (void* x : {S, T, M})
                if (x) (void)ws_free(c, x, stream);
            return Status::error(MISTRA_ERR_NOMEM, std::string("Naive workspace: ") + hipGetErrorString(e));
        }
        // one-pass sums (every source sub-block read once) when the sub-blocks are translates and K has
        // unit stride: S then comes out K-fast ([q][i][k], like T) and the dense GEMM reads both operands
        // K-fast (straight into LDS)
        const bool d2 = p.msum_nsrc == int(p.d * p.d);
        const bool s_k = d2 && p.msum_a && !p.a_mfast && p.kf_dense_k.on, t_one = d2 && p.msum_b && !p.b_nfast;
        for (int64_t z = 0; z < p.nbatch && e == hipSuccess; z++)
        {
            auto multisums = [&](bool isA, hipStream_t st) {
                MultiSumArgs ma;
                std::memset(&ma, 0, sizeof(ma));
                ma.X = isA ? c.A : c.B;
                ma.zoff = isA ? p.h_batch_a[size_t(z)] : p.h_batch_b[size_t(z)];
                ma.rs = isA ? p.a_k : p.b_k;
                ma.cs = isA ? p.a_m : p.b_n;
                ma.rows = p.ks;
                ma.cols = isA ? p.ms : p.ns;
                ma.nsrc = p.msum_nsrc;
                ma.nout = p.nprod;
                ma.coef = isA ? p.msum_coef_a : p.msum_coef_b;
                ma.coef1 = isA ? p.msum_coef1_a : p.msum_coef1_b;
                set_msum_cv(p, isA, ma);
                ma.shift = isA ? p.msum_shift_a : p.msum_shift_b;
                ma.out = isA ? S : T;
                ma.pair = (isA ? p.msum_pair_a : p.msum_pair_b) && ma.zoff % 2 == 0 &&
                          (reinterpret_cast<uintptr_t>(ma.X) & uintptr_t(2 * p.elem - 1)) == 0;
                return launch_operand_multisums(p.elem, ma, st);
            };
            SumArgs sa;
            std::memset(&sa, 0, sizeof(sa));
            sa.X = c.A;
            sa.rs = p.a_m;
            sa.cs = p.a_k;
            sa.rows = p.ms;
            sa.cols = p.ks;
            sa.prods = p.prods;
            sa.nprod = p.nprod;
            sa.which = 0;
            sa.rows_fast = p.a_mfast;
            sa.out = S;
            sa.zoff = p.h_batch_a[size_t(z)];
            // B's sums on the side stream, A's on the caller's, joined before the dense GEMM
            SideFork fk;
            const bool forked = !c.ws && fork_begin(stream, fk);
            const hipStream_t sb = forked ? fk.side : stream;
            e = s_k ? multisums(true, stream) : launch_operand_sums(p.elem, sa, stream);
            sa.X = c.B;
            sa.rs = p.b_k;
            sa.cs = p.b_n;
            sa.rows = p.ks;
            sa.cols = p.ns;
            sa.which = 1;
            sa.rows_fast = !p.b_nfast;
            sa.out = T;
            sa.zoff = p.h_batch_b[size_t(z)];
            const hipError_t eb = e != hipSuccess ? hipSuccess
                                                  : (t_one ? multisums(false, sb) : launch_operand_sums(p.elem, sa, sb));
            if (e == hipSuccess) e = eb;
            if (forked)
            {
                const hipError_t ej = fork_end(stream, fk);
                if (e == hipSuccess) e = ej;
            }
            if (e != hipSuccess) break;

            GemmArgs gd;
            std::memset(&gd, 0, sizeof(gd));
            gd.A = S;
            gd.B = T;
            gd.a_m = s_k ? p.sk_m : p.s_m;
            gd.a_k = s_k ? p.t_k : p.s_k;
            gd.a_krun = s_k ? p.t_krun : p.s_krun;
            gd.b_k = p.t_k;
            gd.b_krun = p.t_krun;
            gd.b_n = p.t_n;
            gd.ms = p.ms;
            gd.ns = p.ns;
            gd.ks = p.ks;
            gd.prods = p.dense_prods;
            gd.nprod = p.nprod;
            gd.nbatch = 1;
            gd.M = M;
            gd.zero = p.zero;
            const KFast& kd = s_k ? p.kf_dense_k : p.kf_dense;
            gd.ka_stride = kd.sa;
            gd.kb_stride = kd.sb;
            timing_begin(stream);
            e = launch_leaf_gemm(p.elem, EPI_STORE, dense_tile(p), 1, kd.mode, gd, !s_k, false, stream);
            timing_end(stream, 2.0 * double(p.nprod) * double(p.ms) * double(p.ns) * double(p.ks));
            if (e != hipSuccess) break;

            CombineArgs ca;
            std::memset(&ca, 0, sizeof(ca));
            ca.C = static_cast<char*>(c.C) + p.h_batch_c[size_t(z)] * p.elem;
            ca.M = M;
            ca.c_m = p.c_m;
            ca.c_n = p.c_n;
            ca.ms = p.ms;
            ca.ns = p.ns;
            ca.nsub = p.nsub;
            ca.sub_roff = p.sub_roff;
            ca.sub_coff = p.sub_coff;
            ca.term_begin = p.term_begin;
            ca.terms = p.terms;
            set_comb_mask(p, ca);
            ca.nprod = p.nprod;
            ca.alpha = c.alpha;
            ca.beta = c.beta;
            ca.nbatch = 1;
            e = launch_combine(p.elem, ca, stream);
        }
        hipError_t f1 = ws_free(c, S, stream), f2 = ws_free(c, T, stream), f3 = ws_free(c, M, stream);
        HIP_TRY(e);
        HIP_TRY(f1);
        HIP_TRY(f2);
        HIP_TRY(f3);
    }

    return run_fringe(p, c, stream);
}

// Classical FP64 / FP32 passes over the remainder of the 2^L split.
Status run_fringe(const Plan& p, const Call& c, hipStream_t stream)
{
    for (const Pass& ps : p.fringe)
    {
        GemmArgs f = base_args(p, c);
        f.ms = ps.ms;
        f.ns = ps.ns;
        f.ks = ps.ks;
        f.prods = (ps.kf.on ? p.prods_fast : p.prods) + ps.prod_index;
        f.nprod = 1;
        f.ka_stride = ps.kf.sa;
        f.kb_stride = ps.kf.sb;
        HIP_TRY(launch_leaf_gemm(p.elem, EPI_MULTI_C, TILE_64, 1, ps.kf.mode, f, p.a_mfast, p.b_nfast, stream));
    }
    return Status();
}

// Low-precision leaf products of an FP64 plan: the Strassen operand sums are formed in FP64 (as for
// Naive), sliced into BF16 digits (Ozaki) or FP8 E4M3 codes with per-128 K-block scales, multiplied on
// BF16 / FP8 MFMA with FP32 accumulation (FP64 across exact chunks / blocks), and the products are
// combined into C in FP64.  The fringe of the 2^L split runs in FP64.
Status run_plan_lowp(const Plan& p, const Call& c, int precision, int nslices, hipStream_t stream)
{
    if (p.elem != 8) return Status::error(MISTRA_ERR_UNSUPPORTED, "low-precision modes take FP64 tensors");
    if (p.general) return Status::error(MISTRA_ERR_UNSUPPORTED, "low-precision modes need a pure contraction");
    if (precision != MISTRA_PREC_FP8_BLOCK && (nslices < 2 || nslices > 7))
        return Status::error(MISTRA_ERR_ARG, "Ozaki slice count must be in 2..7");
    if (precision == MISTRA_PREC_OZAKI_INT8 && nslices > 6)
        return Status::error(MISTRA_ERR_ARG, "INT8 Ozaki slice count must be in 2..6");
    if (p.nprod > 0)
    {
        const bool fp8 = precision == MISTRA_PREC_FP8_BLOCK, i8 = precision == MISTRA_PREC_OZAKI_INT8;
        const int64_t ms = p.ms, ns = p.ns, ks = p.ks, nq = p.nprod;
        auto up = [](int64_t x, int64_t m) { return (x + m - 1) / m * m; };
        const int64_t mp = up(ms, fp8 ? 256 : 128), np = up(ns, fp8 ? 128 : 64), kp = up(ks, fp8 ? 128 : 64);
        const int nsl = fp8 ? 1 : nslices;
        const size_t esz = fp8 || i8 ? 1 : 2;
        struct Buf
        {
            void* p = nullptr;
        } S, T, M, EA, EB, XA, XB;
        const bool direct = !p.core && p.nprod == 1;  // one product = all of C: no staging, no combine
        // Strassen core: the products are staged in M and combined into C in one pass (measured faster
        // than updating the C targets in the GEMM epilogue stage by stage: 70.8 against 65.2 TFLOP/s
        // for INT8 digits at 8192^3, 55.4 against 42.3 at m = n = 16384, k = 1024)
        // (FP8 products are staged in FP32: their E4M3 operands carry about 2^-4 relative error)
        // Operand sums in FP64.  When every term is a translate (kf_core) and an operand's K has unit
        // stride, its sums are written with K fastest by the translate-form kernel (S as [q][i][k], like
        // T), and S's row exponents and digits then use the contiguous-K kernels; a 1-level plan forms
        // such an operand's 7 sums and their digits in one launch instead (no staged sums).
        const bool s_kfast = !fp8 && p.kf_core.on && !p.a_mfast, t_shift = p.kf_core.on && !p.b_nfast;
        const bool ms_a = p.core && p.msum_a && p.msum_nsrc == int(p.d * p.d);
        const bool ms_b = p.core && p.msum_b && p.msum_nsrc == int(p.d * p.d);
        // (level 1: 4 sources, 7 sums; level 2: 16 sources, 49 sums through the Kronecker coefficients)
        const bool fuse1 = p.msum_nsrc == 4 && p.nprod == 7 && p.h_msum_coef_a.size() == 28 &&
                           p.h_msum_coef_b.size() == 28;
        const bool fuse2 = p.msum_nsrc == 16 && p.nprod == 49 && p.h_msum_coef1_a.size() == 28 &&
                           p.h_msum_coef1_b.size() == 28;
        const bool fuse = !fp8 && (fuse1 || fuse2) && !g_dbg_unfused_lowp.load();
        const bool fuse_a = fuse && s_kfast && ms_a, fuse_b = fuse && t_shift && ms_b;
        const size_t nS = fuse_a ? 8 : size_t(nq * ms * ks) * 8, nT = fuse_b ? 8 : size_t(nq * ks * ns) * 8,
                     nM = direct ? 8 : fp8 ? size_t(nq * mp * np) * 4 : size_t(nq * ms * ns) * 8;
        // FP8: one scale byte per row and 128-long K block; digits: one exponent per row
        const size_t nEA = fp8 ? size_t(nq * mp * (kp / 128)) : size_t(nq * ms) * 4;
        const size_t nEB = fp8 ? size_t(nq * np * (kp / 128)) : size_t(nq * ns) * 4;
        const size_t nXA = size_t(nsl) * size_t(nq * mp * kp) * esz, nXB = size_t(nsl) * size_t(nq * np * kp) * esz;
        hipError_t e = hipSuccess;
        for (auto pr : {std::make_pair(&S, nS), std::make_pair(&T, nT), std::make_pair(&M, nM), std::make_pair(&EA, nEA),
                        std::make_pair(&EB, nEB), std::make_pair(&XA, nXA), std::make_pair(&XB, nXB)})
            if (e == hipSuccess) e = stream_alloc(&pr.first->p, pr.second ? pr.second : 1, stream);
        for (int64_t z = 0; z < p.nbatch && e == hipSuccess; z++)
        {
            SumArgs sa;
            std::memset(&sa, 0, sizeof(sa));
            sa.X = c.A;
            sa.prods = p.prods;
            sa.nprod = p.nprod;
            sa.which = 0;
            sa.out = S.p;
            sa.zoff = p.h_batch_a[size_t(z)];
            // one pass over the source sub-blocks when the tables are there (Naive core plans)
            auto multisums = [&](bool isA) {
                MultiSumArgs ma;
                std::memset(&ma, 0, sizeof(ma));
                ma.X = isA ? c.A : c.B;
                ma.zoff = isA ? p.h_batch_a[size_t(z)] : p.h_batch_b[size_t(z)];
                ma.rs = isA ? p.a_k : p.b_k;
                ma.cs = isA ? p.a_m : p.b_n;
                ma.rows = ks;
                ma.cols = isA ? ms : ns;
                ma.nsrc = p.msum_nsrc;
                ma.nout = p.nprod;
                ma.coef = isA ? p.msum_coef_a : p.msum_coef_b;
                ma.coef1 = isA ? p.msum_coef1_a : p.msum_coef1_b;
                set_msum_cv(p, isA, ma);
                ma.shift = isA ? p.msum_shift_a : p.msum_shift_b;
                ma.out = isA ? S.p : T.p;
                ma.pair = (isA ? p.msum_pair_a : p.msum_pair_b) && ma.zoff % 2 == 0 &&
                          (reinterpret_cast<uintptr_t>(ma.X) & uintptr_t(15)) == 0;
                return launch_operand_multisums(8, ma, stream);
            };
            if (s_kfast)
            {
                sa.rs = p.a_k;
                sa.cs = p.a_m;
                sa.rows = ks;
                sa.cols = ms;
                sa.rows_fast = true;
                sa.shifted = true;
                sa.prods = p.prods_fast;
            }
            else
            {
                sa.rs = p.a_m;
                sa.cs = p.a_k;
                sa.rows = ms;
                sa.cols = ks;
                sa.rows_fast = p.a_mfast;
            }
            auto sum_slices = [&](bool isA) {
                SumSliceArgs fa;
                std::memset(&fa, 0, sizeof(fa));
                fa.X = static_cast<const double*>(isA ? c.A : c.B);
                fa.zoff = isA ? p.h_batch_a[size_t(z)] : p.h_batch_b[size_t(z)];
                fa.rs = isA ? p.a_k : p.b_k;
                fa.rs_step = isA ? p.a_k_step : p.b_k_step;
                fa.rs0 = fa.rs_step >= 0 ? (isA ? p.h_a_k[0] : p.h_b_k[0]) : 0;
                fa.cs = isA ? p.a_m : p.b_n;
                fa.ks = ks;
                fa.rows = isA ? ms : ns;
                fa.cp = isA ? mp : np;
                fa.kp = kp;
                fa.level = fuse2 ? 2 : 1;
                const std::vector<int64_t>& sh = isA ? p.h_msum_shift_a : p.h_msum_shift_b;
                const std::vector<double>& cf = fuse2 ? (isA ? p.h_msum_coef1_a : p.h_msum_coef1_b)
                                                      : (isA ? p.h_msum_coef_a : p.h_msum_coef_b);
                for (int i = 0; i < p.msum_nsrc; i++) fa.shift[i] = sh[size_t(i)];
                for (int i = 0; i < 28; i++) fa.coef[i] = cf[size_t(i)];
                fa.e = static_cast<int*>(isA ? EA.p : EB.p);
                fa.out = isA ? XA.p : XB.p;
                return launch_ozaki_sum_slices(nsl, i8, fa, stream);
            };
            e = fuse_a ? sum_slices(true) : s_kfast && ms_a ? multisums(true) : launch_operand_sums(8, sa, stream);
            if (e != hipSuccess) break;
            sa.X = c.B;
            sa.rs = p.b_k;
            sa.cs = p.b_n;
            sa.rows = ks;
            sa.cols = ns;
            sa.which = 1;
            sa.rows_fast = !p.b_nfast;
            sa.shifted = t_shift;
            sa.prods = t_shift ? p.prods_fast : p.prods;
            sa.out = T.p;
            sa.zoff = p.h_batch_b[size_t(z)];
            e = fuse_b ? sum_slices(false) : t_shift && ms_b ? multisums(false) : launch_operand_sums(8, sa, stream);
            if (e != hipSuccess) break;

            LowpSliceArgs la;
            la.S = static_cast<const double*>(S.p);
            la.T = static_cast<const double*>(T.p);
            la.ms = ms;
            la.ns = ns;
            la.ks = ks;
            la.mp = mp;
            la.np = np;
            la.kp = kp;
            la.nprod = p.nprod;
            la.ea = static_cast<int*>(EA.p);
            la.eb = static_cast<int*>(EB.p);
            la.outA = XA.p;
            la.outB = XB.p;
            la.s_kfast = s_kfast;
            la.skip_a = fuse_a;
            la.skip_b = fuse_b;
            LowpGemmArgs lg{ms, ns, mp, np, kp, p.nprod, nullptr, nullptr, nullptr, c.alpha, c.beta, 0, p.nprod,
                            g_dbg_i8_fold.load(), fp8 && !direct};
            if (direct)
            {
                lg.C = reinterpret_cast<double*>(static_cast<char*>(c.C) + p.h_batch_c[size_t(z)] * p.elem);
                lg.c_m = p.c_m;
                lg.c_n = p.c_n;
            }
            const double pairs = fp8 ? 1.0 : double(nsl * (nsl + 1) / 2);
            auto gemm = [&](const LowpGemmArgs& g) {
                timing_begin(stream);
                hipError_t r;
                if (fp8)
                    r = launch_fp8_gemm(XA.p, XB.p, static_cast<const uint32_t*>(EA.p), static_cast<const uint32_t*>(EB.p),
                                        static_cast<double*>(M.p), g, stream);
                else if (i8)
                    r = launch_ozaki_i8_gemm(nsl, XA.p, XB.p, la.ea, la.eb, static_cast<double*>(M.p), g, stream);
                else
                    r = launch_ozaki_gemm(nsl, XA.p, XB.p, la.ea, la.eb, static_cast<double*>(M.p), g, stream);
                timing_end(stream, 2.0 * double(g.nlaunch) * double(mp) * double(np) * double(kp) * pairs);
                return r;
            };
            if (fp8)
                e = launch_fp8_slices(la, static_cast<uint32_t*>(EA.p), static_cast<uint32_t*>(EB.p), stream);
            else
            {
                e = launch_lowp_exponents(la, stream);
                if (e == hipSuccess) e = i8 ? launch_ozaki_i8_slices(nsl, la, stream) : launch_ozaki_slices(nsl, la, stream);
            }
            if (e != hipSuccess) break;
            e = gemm(lg);
            if (e != hipSuccess) break;
            if (direct) continue;

            CombineArgs ca;
            std::memset(&ca, 0, sizeof(ca));
            ca.C = static_cast<char*>(c.C) + p.h_batch_c[size_t(z)] * p.elem;
            ca.M = M.p;
            ca.c_m = p.c_m;
            ca.c_n = p.c_n;
            ca.ms = ms;
            ca.ns = ns;
            ca.nsub = p.nsub;
            ca.sub_roff = p.sub_roff;
            ca.sub_coff = p.sub_coff;
            ca.term_begin = p.term_begin;
            ca.terms = p.terms;
            set_comb_mask(p, ca);
            ca.nprod = p.nprod;
            ca.alpha = c.alpha;
            ca.beta = c.beta;
            ca.nbatch = 1;
            ca.m_elem = fp8 ? 4 : 0;
            ca.m_tiled = fp8 ? 1 : 0;  // (the FP8 GEMM's tiles)
            ca.m_mp = mp;
            ca.m_np = np;
            e = launch_combine(8, ca, stream);
        }
        for (Buf* b : {&S, &T, &M, &EA, &EB, &XA, &XB})
            if (b->p) (void)hipFreeAsync(b->p, stream);
        HIP_TRY(e);
    }
    return run_fringe(p, c, stream);
}

}  // namespace mistra
