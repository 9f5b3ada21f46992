// Host-side index layer and Strassen tables (see plan.hpp for the reference map).
#include "plan.hpp"

#include <algorithm>
#include <numeric>
#include <sstream>

namespace mistra {

void diagonal(unsigned ndim, const len_type* len, const stride_type* stride, const label_type* idx,
              std::vector<len_type>& len_out, std::vector<stride_type>& stride_out,
              std::vector<label_type>& idx_out)
{
    // src/util/tensor.hpp:305-342: visit dims in label order; a repeated label adds its stride to the
    // kept dim (a diagonal), a length-1 dim is dropped.
    std::vector<unsigned> order(ndim);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(),
                     [&](unsigned a, unsigned b) { return idx[a] < idx[b]; });
    len_out.clear();
    stride_out.clear();
    idx_out.clear();
    for (unsigned i = 0; i < ndim; i++)
    {
        unsigned d = order[i];
        bool repeat = i > 0 && idx[d] == idx[order[i - 1]];
        if (len[d] == 1) continue;
        if (repeat && !idx_out.empty() && idx_out.back() == idx[d])
            stride_out.back() += stride[d];
        else
        {
            len_out.push_back(len[d]);
            stride_out.push_back(stride[d]);
            idx_out.push_back(idx[d]);
        }
    }
}

namespace {

struct Dims
{
    std::vector<len_type> len;
    std::vector<stride_type> stride;
    std::vector<label_type> idx;
    int find(label_type l) const
    {
        for (size_t i = 0; i < idx.size(); i++)
            if (idx[i] == l) return int(i);
        return -1;
    }
};

}  // namespace

std::vector<unsigned> sort_by_stride(const std::vector<stride_type>& s1, const std::vector<stride_type>& s2)
{
    std::vector<unsigned> order(s1.size());
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](unsigned a, unsigned b) {
        if (s1[a] != s1[b]) return s1[a] < s1[b];
        if (!s2.empty() && s2[a] != s2[b]) return s2[a] < s2[b];
        return false;
    });
    return order;
}

void fold(IndexGroup& g, int nstrides)
{
    // src/util/tensor.hpp:270-303: order by the first stride vector, then merge index i into the
    // previous kept one when stride[i] == stride[i-1] * len[i-1] in every stride vector.
    std::vector<std::vector<stride_type>*> s = {&g.sA, &g.sB, &g.sC};
    std::vector<std::vector<stride_type>*> used;
    for (auto* v : s)
        if (!v->empty() || g.len.empty()) used.push_back(v);
    if (g.len.empty()) return;
    (void)nstrides;
    std::vector<stride_type> empty;
    std::vector<unsigned> order = sort_by_stride(*used[0], empty);

    IndexGroup out;
    for (size_t i = 0; i < order.size(); i++)
    {
        unsigned d = order[i];
        bool contiguous = false;
        if (i > 0)
        {
            unsigned p = order[i - 1];
            contiguous = true;
            for (auto* v : used)
                if ((*v)[d] != (*v)[p] * g.len[p]) contiguous = false;
        }
        if (contiguous)
            out.len.back() *= g.len[d];
        else
        {
            out.idx.push_back(g.idx[d]);
            out.len.push_back(g.len[d]);
            if (!g.sA.empty()) out.sA.push_back(g.sA[d]);
            if (!g.sB.empty()) out.sB.push_back(g.sB[d]);
            if (!g.sC.empty()) out.sC.push_back(g.sC[d]);
        }
    }
    g = std::move(out);
}

bool group_indices(const tblis::tblis_tensor& A, const label_type* idx_A,
                   const tblis::tblis_tensor& B, const label_type* idx_B,
                   const tblis::tblis_tensor& C, const label_type* idx_C,
                   Grouping& g, std::string& err)
{
    Dims a, b, c;
    diagonal(A.ndim, A.len, A.stride, idx_A, a.len, a.stride, a.idx);
    diagonal(B.ndim, B.len, B.stride, idx_B, b.len, b.stride, b.idx);
    diagonal(C.ndim, C.len, C.stride, idx_C, c.len, c.stride, c.idx);

    // src/stra_iface/3t/stra_mult.cxx:89-125: every label lands in exactly one of 7 groups, listed
    // in label order (the set operations of stl_ext sort their inputs).
    std::vector<label_type> labels;
    for (auto* d : {&a, &b, &c})
        for (label_type l : d->idx) labels.push_back(l);
    std::sort(labels.begin(), labels.end());
    labels.erase(std::unique(labels.begin(), labels.end()), labels.end());

    g = Grouping();
    for (label_type l : labels)
    {
        int ia = a.find(l), ib = b.find(l), ic = c.find(l);
        len_type la = ia >= 0 ? a.len[ia] : -1, lb = ib >= 0 ? b.len[ib] : -1, lc = ic >= 0 ? c.len[ic] : -1;
        len_type L = std::max(la, std::max(lb, lc));
        if ((la >= 0 && la != L) || (lb >= 0 && lb != L) || (lc >= 0 && lc != L))
        {
            std::ostringstream os;
            os << "index '" << l << "' has mismatched lengths (" << la << ", " << lb << ", " << lc << ")";
            err = os.str();
            return false;
        }
        IndexGroup* grp;
        if (ia >= 0 && ib >= 0 && ic >= 0) grp = &g.ABC;
        else if (ia >= 0 && ib >= 0) grp = &g.AB;
        else if (ia >= 0 && ic >= 0) grp = &g.AC;
        else if (ib >= 0 && ic >= 0) grp = &g.BC;
        else if (ia >= 0) grp = &g.A_only;
        else if (ib >= 0) grp = &g.B_only;
        else grp = &g.C_only;
        grp->idx.push_back(l);
        grp->len.push_back(L);
        if (ia >= 0) grp->sA.push_back(a.stride[ia]);
        if (ib >= 0) grp->sB.push_back(b.stride[ib]);
        if (ic >= 0) grp->sC.push_back(c.stride[ic]);
    }
    // src/stra_iface/3t/stra_mult.cxx:174-180: fold each group, ordered by its first stride vector.
    fold(g.ABC, 3);
    fold(g.AB, 2);
    fold(g.AC, 2);
    fold(g.BC, 2);
    fold(g.A_only, 1);
    fold(g.B_only, 1);
    fold(g.C_only, 1);
    return true;
}

std::vector<int64_t> scatter(const std::vector<len_type>& len, const std::vector<stride_type>& stride)
{
    int64_t n = 1;
    for (len_type l : len) n *= l;
    std::vector<int64_t> out(size_t(n > 0 ? n : 0));
    if (n <= 0) return out;
    size_t nd = len.size();
    std::vector<int64_t> digit(nd, 0);
    int64_t off = 0;
    for (int64_t p = 0; p < n; p++)
    {
        out[size_t(p)] = off;
        // increment the mixed-radix counter, first index fastest
        for (size_t d = 0; d < nd; d++)
        {
            if (++digit[d] < len[d]) { off += stride[d]; break; }
            off -= stride[d] * (len[d] - 1);
            digit[d] = 0;
        }
    }
    return out;
}

std::vector<int64_t> run_strides(const std::vector<int64_t>& scat, int run)
{
    const int64_t n = int64_t(scat.size());
    std::vector<int64_t> out(scat.size(), 0);
    if (n == 0) return out;
    // diff[p] = scat[p+1] - scat[p]; a run starting at p is affine iff the diffs p..p+run-2 agree.
    // Walk backwards keeping the length of the constant-diff streak that starts at p.
    std::vector<int64_t> streak(size_t(n), 0);
    for (int64_t p = n - 2; p >= 0; p--)
    {
        int64_t d = scat[size_t(p + 1)] - scat[size_t(p)];
        bool same = p + 2 < n && scat[size_t(p + 2)] - scat[size_t(p + 1)] == d;
        streak[size_t(p)] = same ? streak[size_t(p + 1)] + 1 : 1;
    }
    for (int64_t p = 0; p < n; p++)
    {
        int64_t need = std::min<int64_t>(run, n - p) - 1;  // number of diffs inside the run
        if (need <= 0) { out[size_t(p)] = 1; continue; }    // single element: any stride works
        int64_t d = scat[size_t(p + 1)] - scat[size_t(p)];
        if (streak[size_t(p)] >= need && d != 0) out[size_t(p)] = d;
    }
    return out;
}

// ---- Strassen tables ---------------------------------------------------------------------------------

namespace {

// The 7 products of src/stra_internal/3t/stra_mult.cxx:426-591 (U, V, W of src/stra_internal/3t/strassen),
// sub-block ids row-major in a 2x2 split: 0 = (0,0), 1 = (0,1), 2 = (1,0), 3 = (1,1).
std::vector<Product> make_1level()
{
    auto P = [](std::vector<SubTerm> a, std::vector<SubTerm> b, std::vector<SubTerm> c) {
        return Product{std::move(a), std::move(b), std::move(c)};
    };
    return {
        P({{0, 1}, {3, 1}}, {{0, 1}, {3, 1}}, {{0, 1}, {3, 1}}),   // M0 = (A0+A3)(B0+B3)
        P({{2, 1}, {3, 1}}, {{0, 1}}, {{2, 1}, {3, -1}}),           // M1 = (A2+A3) B0
        P({{0, 1}}, {{1, 1}, {3, -1}}, {{1, 1}, {3, 1}}),           // M2 = A0 (B1-B3)
        P({{3, 1}}, {{0, -1}, {2, 1}}, {{0, 1}, {2, 1}}),           // M3 = A3 (B2-B0)
        P({{0, 1}, {1, 1}}, {{3, 1}}, {{0, -1}, {1, 1}}),           // M4 = (A0+A1) B3
        P({{0, -1}, {2, 1}}, {{0, 1}, {1, 1}}, {{3, 1}}),           // M5 = (A2-A0)(B0+B1)
        P({{1, 1}, {3, -1}}, {{2, 1}, {3, 1}}, {{0, 1}}),           // M6 = (A1-A3)(B2+B3)
    };
}

// Kronecker square of the 1-level table: outer block (s1,t1), inner (s2,t2) -> row 2*s1+s2,
// col 2*t1+t2 of a 4x4 split, sub-block id = 4*row + col.
std::vector<SubTerm> kron(const std::vector<SubTerm>& outer, const std::vector<SubTerm>& inner)
{
    std::vector<SubTerm> out;
    for (const SubTerm& o : outer)
        for (const SubTerm& i : inner)
        {
            int s1 = o.subid / 2, t1 = o.subid % 2, s2 = i.subid / 2, t2 = i.subid % 2;
            out.push_back({4 * (2 * s1 + s2) + (2 * t1 + t2), o.coeff * i.coeff});
        }
    return out;
}

std::vector<Product> make_2level()
{
    const auto& one = strassen_products_1level();
    std::vector<Product> out;
    for (const Product& po : one)
        for (const Product& pi : one)
            out.push_back({kron(po.A, pi.A), kron(po.B, pi.B), kron(po.C, pi.C)});
    return out;
}

}  // namespace

const std::vector<Product>& strassen_products_1level()
{
    static const std::vector<Product> t = make_1level();
    return t;
}

const std::vector<Product>& strassen_products_2level()
{
    static const std::vector<Product> t = make_2level();
    return t;
}

const std::vector<Product>& products_for_level(int level)
{
    static const std::vector<Product> l0 = {Product{{{0, 1}}, {{0, 1}}, {{0, 1}}}};
    if (level == 1) return strassen_products_1level();
    if (level == 2) return strassen_products_2level();
    return l0;
}

}  // namespace mistra
