// Batched contractions (mistra_batch_mult).
//
// Default: one persistent launch (batch_persist.hip).  Every device-resident ABC contraction contributes
// its leaf products (Strassen stages, then fringe passes) as items with a round number; each item's
// operand sums are packed into dense K-fast panels by pack jobs, and its tiles are GEMM jobs of one
// kernel variant; the launch works through all of them from one device queue.  Tables, panels and the
// dependency counters are built once per set of entries and cached; MISTRA_FLAG_GRAPH records the
// launch in a HIP graph (one kernel node: the counters reset themselves at the end of every launch).
//
// MISTRA_FLAG_BATCH_GROUPED: the jobs of one round and one kernel variant become a single grouped launch
// (leaf_gemm_grouped: a 1-D grid whose workgroups look their job up in a device table); rounds run in
// stream order, which keeps every contraction's stage order; MISTRA_FLAG_GRAPH captures that launch
// sequence.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <list>
#include <map>
#include <mutex>
#include <sstream>
#include <tuple>

#include "batch_persist.hpp"
#include "driver.hpp"

namespace mistra {

std::atomic<long long> g_dbg_batch_spin{0};
std::atomic<unsigned> g_dbg_batch_bias{0};

double scalar_value(const tblis::tblis_tensor* t)
{
    if (t->type == tblis::TYPE_DOUBLE)
    {
        double v;
        std::memcpy(&v, t->scalar, sizeof(v));
        return v;
    }
    float f;
    std::memcpy(&f, t->scalar, sizeof(f));
    return double(f);
}

namespace {

bool on_device(const void* p)
{
    hipPointerAttribute_t a;
    if (!p || hipPointerGetAttributes(&a, p) != hipSuccess)
    {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice;
}

struct Group
{
    int elem, tile, terms, kmode;
    bool a_mfast, b_nfast;
    size_t job_off, first_off;  // offsets of the job and first tables in the device buffer
    int njobs;
    int64_t blocks;
    double flops;
};

struct BatchGraph
{
    std::string key;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void* tables = nullptr;
    std::vector<std::shared_ptr<Plan>> plans;
    ~BatchGraph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (tables) (void)hipFree(tables);
    }
};

std::mutex g_graph_mutex;
std::list<std::shared_ptr<BatchGraph>> g_graphs;
constexpr size_t kMaxGraphs = 16;

void append_tensor(std::ostringstream& k, const tblis::tblis_tensor* t, const label_type* idx)
{
    k << int(t->type) << '@' << t->data << '*' << scalar_value(t) << '[';
    for (unsigned i = 0; i < t->ndim; i++) k << idx[i] << ':' << t->len[i] << ':' << t->stride[i] << ',';
    k << ']';
}

// Plans and launch groups of a batch; `singles` lists the entries that run as individual calls.
Status prepare(int count, const mistra_contraction* e, std::vector<std::shared_ptr<Plan>>& plans,
               std::vector<int>& singles, std::vector<GemmArgs>& jobs_host, std::vector<int64_t>& first_host,
               std::vector<Group>& groups, hipStream_t stream)
{
    std::vector<LaunchJob> all;
    for (int i = 0; i < count; i++)
    {
        const mistra_contraction& x = e[i];
        if (!x.A || !x.B || !x.C) return Status::error(MISTRA_ERR_ARG, "null tensor in batch entry");
        int level = x.level, impl = x.impl;
        if (level < 0 && impl >= 0)
            level = best_level(x.A, x.idx_A, x.B, x.idx_B, x.C, x.idx_C, impl);  // the model's level for that impl
        else if (level < 0 || impl < 0)
        {
            TuneChoice ch;
            Status st = autotune(x.A, x.idx_A, x.B, x.idx_B, x.C, x.idx_C, false, stream, ch);
            if (!st.ok()) return st;
            if (level < 0) level = ch.level;
            if (impl < 0) impl = ch.impl;
        }
        const double alpha = scalar_value(x.A) * scalar_value(x.B), beta = scalar_value(x.C);
        bool empty = false;
        for (const tblis::tblis_tensor* t : {x.A, x.B, static_cast<const tblis::tblis_tensor*>(x.C)})
            for (unsigned d = 0; d < t->ndim; d++)
                if (t->len[d] == 0) empty = true;
        if (impl != IMPL_ABC || alpha == 0.0 || empty || !on_device(x.A->data) || !on_device(x.B->data) ||
            !on_device(x.C->data))
        {
            singles.push_back(i);
            continue;
        }
        std::shared_ptr<Plan> plan;
        // fused sums: every launch of the plan is a leaf launch, so it can join a grouped launch
        Status st = get_plan(x.A, x.idx_A, x.B, x.idx_B, x.C, x.idx_C, level, IMPL_ABC, -1, MISTRA_FLAG_FUSED_SUMS,
                             plan);
        if (!st.ok()) return st;
        if (plan->general || plan->outer_staged)
        {
            singles.push_back(i);
            continue;
        }
        Call call{x.A->data, x.B->data, x.C->data, alpha, beta};
        if (plan->transposed) std::swap(call.A, call.B);
        st = plan_abc_jobs(*plan, call, all);
        if (!st.ok()) return st;
        plans.push_back(plan);
    }
    // group by (round, variant)
    std::map<std::tuple<int, int, int, int, int, bool, bool>, std::vector<const LaunchJob*>> by;
    for (const LaunchJob& j : all)
        by[std::make_tuple(j.round, j.elem, j.tile, j.terms, j.kmode, j.a_mfast, j.b_nfast)].push_back(&j);
    for (auto& kv : by)
    {
        Group gr;
        std::tie(std::ignore, gr.elem, gr.tile, gr.terms, gr.kmode, gr.a_mfast, gr.b_nfast) = kv.first;
        gr.job_off = jobs_host.size();
        gr.first_off = first_host.size();
        gr.njobs = int(kv.second.size());
        gr.blocks = 0;
        gr.flops = 0;
        for (const LaunchJob* j : kv.second)
        {
            jobs_host.push_back(j->g);
            first_host.push_back(gr.blocks);
            gr.blocks += int64_t(j->g.tiles_m) * j->g.tiles_n * j->g.nprod * j->g.nbatch;
            gr.flops += j->flops;
        }
        first_host.push_back(gr.blocks);
        groups.push_back(gr);
    }
    return Status();
}

Status upload_tables(const std::vector<GemmArgs>& jobs, const std::vector<int64_t>& first, void*& buf,
                     size_t& first_base, hipStream_t stream, bool async)
{
    const size_t jb = jobs.size() * sizeof(GemmArgs);
    first_base = (jb + 255) / 256 * 256;
    const size_t total = first_base + first.size() * sizeof(int64_t);
    std::vector<unsigned char> host(total, 0);
    if (jb) std::memcpy(host.data(), jobs.data(), jb);
    if (!first.empty()) std::memcpy(host.data() + first_base, first.data(), first.size() * sizeof(int64_t));
    hipError_t e = async ? stream_alloc(&buf, total ? total : 1, stream) : hipMalloc(&buf, total ? total : 1);
    if (e != hipSuccess) return Status::error(MISTRA_ERR_NOMEM, hipGetErrorString(e));
    e = async ? hipMemcpyAsync(buf, host.data(), total, hipMemcpyHostToDevice, stream)
              : hipMemcpy(buf, host.data(), total, hipMemcpyHostToDevice);
    if (e == hipSuccess && async) e = hipStreamSynchronize(stream);  // the host staging buffer dies here
    if (e != hipSuccess) return Status::error(MISTRA_ERR_HIP, hipGetErrorString(e));
    return Status();
}

Status launch_groups(const std::vector<Group>& groups, void* tables, size_t first_base, hipStream_t stream,
                     bool timed)
{
    for (const Group& gr : groups)
    {
        Grouped q;
        q.jobs = reinterpret_cast<const GemmArgs*>(static_cast<char*>(tables) + gr.job_off * sizeof(GemmArgs));
        q.first = reinterpret_cast<const int64_t*>(static_cast<char*>(tables) + first_base +
                                                   gr.first_off * sizeof(int64_t));
        q.njobs = gr.njobs;
        q.blocks = gr.blocks;
        if (timed) timing_begin(stream);
        hipError_t e = launch_leaf_grouped(gr.elem, gr.tile, gr.terms, gr.kmode, q, gr.a_mfast, gr.b_nfast, stream);
        if (timed) timing_end(stream, gr.flops);
        if (e != hipSuccess) return Status::error(MISTRA_ERR_HIP, std::string("grouped launch: ") + hipGetErrorString(e));
    }
    return Status();
}

Status batch_grouped(int count, const mistra_contraction* entries, unsigned flags, hipStream_t stream);

// ---- persistent launch ------------------------------------------------------------------------------

constexpr long long kDefaultSpin = 1ll << 22;  // polls before a wait gives up (each a coherent load and a short sleep: seconds in all)

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }
int elem_tk(int elem) { return elem == 8 ? 16 : 32; }  // K-tile of the leaf kernel: 128-byte K rows

struct PersistBatch
{
    std::string key;
    std::vector<std::shared_ptr<Plan>> plans;
    std::vector<int> singles;
    void* tables = nullptr;
    void* work = nullptr;
    PersistArgs args;
    int elem = 8, tile = TILE_128, grid = 0;
    double flops = 0.0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~PersistBatch()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (tables) (void)hipFree(tables);
        if (work) (void)hipFree(work);
    }
};

std::mutex g_persist_mutex;
std::list<std::shared_ptr<PersistBatch>> g_persist;  // most recent first
constexpr size_t kMaxPersist = 4;

// One leaf product (or fringe pass) of one contraction, before the device addresses are known.
struct HostItem
{
    int entry, round;
    int64_t ms, ns, ks, kpad;
    ProdDesc prod;           // the plan's product: A, B terms (sub-block offsets) and C targets
    const Plan* plan;
    Call call;
    double alpha, beta;
};

// Replays the launch's queue discipline on the host tables, segment by segment in queue order (a job
// waits only for jobs before it, so this is the order every wait must already be met in): the segments
// tile the queue without gaps, every panel and every product appears exactly once, every pack chunk
// advances its product's counter, every product's pack and round targets are reached when its first
// tile's turn comes, and at the end every product's counter equals its target.
Status check_queue(const Segment* segs, int nsegs, const PackDesc* packs, int npacks, const TileItem* items, int nitems,
                   int nentries, int64_t njobs)
{
    std::vector<int64_t> packed(size_t(nitems), 0), done(size_t(nentries), 0);
    std::vector<char> seen_pack(size_t(npacks), 0), seen_item(size_t(nitems), 0);
    int64_t pos = 0;
    for (int s = 0; s < nsegs; s++)
    {
        const Segment& S = segs[s];
        const std::string where = "queue: segment " + std::to_string(s) + " at " + std::to_string(pos);
        if (S.first_job != pos) return Status::error(MISTRA_ERR_ARG, where + ": segments leave a gap or overlap");
        if (S.kind == 0)
        {
            if (S.index < 0 || S.index >= npacks || seen_pack[size_t(S.index)]++)
                return Status::error(MISTRA_ERR_ARG, where + ": panel index out of range or repeated");
            const PackDesc& P = packs[S.index];
            if (P.first_job != pos || P.item < 0 || P.item >= nitems)
                return Status::error(MISTRA_ERR_ARG, where + ": panel position or product out of step");
            const int64_t n = pack_jobs_of(P.rows, P.ld);
            packed[size_t(P.item)] += n;
            pos += n;
            continue;
        }
        if (S.kind != 1 || S.index < 0 || S.index >= nitems || seen_item[size_t(S.index)]++)
            return Status::error(MISTRA_ERR_ARG, where + ": product index out of range or repeated");
        const TileItem& it = items[S.index];
        if (it.first_job != pos) return Status::error(MISTRA_ERR_ARG, where + ": product position out of step");
        if (packed[size_t(S.index)] < it.need_pack)
            return Status::error(MISTRA_ERR_ARG, where + ": product " + std::to_string(S.index) + " waits for " +
                                                     std::to_string(it.need_pack) + " pack chunks, " +
                                                     std::to_string(packed[size_t(S.index)]) + " queued before it");
        if (it.entry < 0 || it.entry >= nentries || done[size_t(it.entry)] < it.need_entry)
            return Status::error(MISTRA_ERR_ARG, where + ": product waits for " + std::to_string(it.need_entry) +
                                                     " earlier tiles of contraction " + std::to_string(it.entry));
        const int64_t n = int64_t(it.g.tiles_m) * it.g.tiles_n;
        done[size_t(it.entry)] += n;
        pos += n;
    }
    if (pos != njobs) return Status::error(MISTRA_ERR_ARG, "queue: segments cover " + std::to_string(pos) + " of " +
                                                              std::to_string(njobs) + " positions");
    for (int i = 0; i < npacks; i++)
        if (!seen_pack[size_t(i)]) return Status::error(MISTRA_ERR_ARG, "queue: panel " + std::to_string(i) + " never packed");
    for (int i = 0; i < nitems; i++)
    {
        if (!seen_item[size_t(i)]) return Status::error(MISTRA_ERR_ARG, "queue: product " + std::to_string(i) + " never queued");
        if (packed[size_t(i)] != items[i].need_pack)
            return Status::error(MISTRA_ERR_ARG, "queue: product " + std::to_string(i) + " packs " +
                                                     std::to_string(packed[size_t(i)]) + " chunks, expects " +
                                                     std::to_string(items[i].need_pack));
    }
    return Status();
}

// host_only: plans built on the host alone (no device tables), no device checks or allocations; the tables
// are laid out at placeholder addresses and only checked with check_queue.
Status prepare_persistent(int count, const mistra_contraction* e, hipStream_t stream, PersistBatch& pb,
                          bool host_only = false)
{
    std::vector<HostItem> items;
    int elem = 0, nentries = 0;
    double flops = 0.0;
    for (int i = 0; i < count; i++)
    {
        const mistra_contraction& x = e[i];
        if (!x.A || !x.B || !x.C) return Status::error(MISTRA_ERR_ARG, "null tensor in batch entry");
        int level = x.level, impl = x.impl;
        if (level < 0 && impl >= 0)
            level = best_level(x.A, x.idx_A, x.B, x.idx_B, x.C, x.idx_C, impl);
        else if (level < 0 || impl < 0)
        {
            TuneChoice ch;
            Status st = autotune(x.A, x.idx_A, x.B, x.idx_B, x.C, x.idx_C, false, stream, ch);
            if (!st.ok()) return st;
            if (level < 0) level = ch.level;
            if (impl < 0) impl = ch.impl;
        }
        const double alpha = scalar_value(x.A) * scalar_value(x.B), beta = scalar_value(x.C);
        bool empty = false;
        for (const tblis::tblis_tensor* t : {x.A, x.B, static_cast<const tblis::tblis_tensor*>(x.C)})
            for (unsigned d = 0; d < t->ndim; d++)
                if (t->len[d] == 0) empty = true;
        const int xelem = x.C->type == tblis::TYPE_DOUBLE ? 8 : (x.C->type == tblis::TYPE_FLOAT ? 4 : 0);
        const bool dev_ok = host_only || (on_device(x.A->data) && on_device(x.B->data) && on_device(x.C->data));
        if (impl != IMPL_ABC || alpha == 0.0 || empty || !xelem || (elem && xelem != elem) ||
            x.A->type != x.C->type || x.B->type != x.C->type || !dev_ok || level > 2)
        {
            pb.singles.push_back(i);
            continue;
        }
        std::shared_ptr<Plan> plan;
        Status st;
        if (host_only)
        {
            plan = std::make_shared<Plan>();
            st = build_plan_host(x.A, x.idx_A, x.B, x.idx_B, x.C, x.idx_C, level, IMPL_ABC, -1, MISTRA_FLAG_FUSED_SUMS,
                                 *plan);
            if (st.ok() && !plan->general)
            {
                std::vector<int> order;
                make_stages(*plan, order);
                plan->h_stage_order = order;
            }
        }
        else
            st = get_plan(x.A, x.idx_A, x.B, x.idx_B, x.C, x.idx_C, level, IMPL_ABC, -1, MISTRA_FLAG_FUSED_SUMS, plan);
        if (!st.ok()) return st;
        if (plan->general || plan->nbatch != 1 || plan->nprod == 0)
        {
            pb.singles.push_back(i);
            continue;
        }
        const Plan& p = *plan;
        const int TK = elem_tk(xelem);
        auto kpad_of = [&](int64_t ks) { return (ks + TK - 1) / TK * TK; };
        // every panel's element offsets must fit the 32-bit offsets of the leaf kernel's loads
        const int64_t lim = (int64_t(1) << 31) / xelem;
        bool fits = true;
        auto check = [&](int64_t ms, int64_t ns, int64_t ks) {
            fits = fits && ms * kpad_of(ks) < lim && ns * kpad_of(ks) < lim;
        };
        check(p.ms, p.ns, p.ks);
        for (const Pass& ps : p.fringe) check(ps.ms, ps.ns, ps.ks);
        if (!fits)
        {
            pb.singles.push_back(i);
            continue;
        }
        elem = xelem;
        Call call{x.A->data, x.B->data, x.C->data, alpha, beta};
        if (p.transposed) std::swap(call.A, call.B);
        call.A = static_cast<const char*>(call.A) + p.h_batch_a[0] * elem;
        call.B = static_cast<const char*>(call.B) + p.h_batch_b[0] * elem;
        call.C = static_cast<char*>(call.C) + p.h_batch_c[0] * elem;
        const int entry = nentries++;
        const int nst = int(p.h_stage_begin.size()) - 1;
        for (int st_i = 0; st_i < nst; st_i++)
            for (int j = p.h_stage_begin[size_t(st_i)]; j < p.h_stage_begin[size_t(st_i) + 1]; j++)
            {
                HostItem it{entry, st_i, p.ms, p.ns, p.ks, kpad_of(p.ks), p.h_prods[size_t(p.h_stage_order[size_t(j)])],
                            &p, call, alpha, beta};
                it.prod.first_mask = p.h_stage_first[size_t(j)];
                items.push_back(it);
                flops += 2.0 * double(p.ms) * double(p.ns) * double(p.ks);
            }
        int round = nst;
        for (const Pass& ps : p.fringe)
        {
            if (ps.ms <= 0 || ps.ns <= 0 || ps.ks <= 0) continue;
            items.push_back(HostItem{entry, round++, ps.ms, ps.ns, ps.ks, kpad_of(ps.ks), p.h_prods[size_t(ps.prod_index)],
                                     &p, call, alpha, beta});
            flops += 2.0 * double(ps.ms) * double(ps.ns) * double(ps.ks);
        }
        pb.plans.push_back(plan);
    }
    pb.elem = elem ? elem : 8;
    pb.flops = flops;
    if (items.empty()) return Status();
    int dev = 0, cus = 256;
    if (!host_only &&
        (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess))
        return Status::error(MISTRA_ERR_HIP, "device attributes");
    // 128 x 128 tiles when the batch has enough of them to fill the CUs twice, else 64 x 64
    int64_t t128 = 0;
    for (const HostItem& it : items) t128 += ((it.ms + 127) / 128) * ((it.ns + 127) / 128);
    pb.tile = t128 >= 2 * cus ? TILE_128 : TILE_64;
    if (tile_override() == TILE_128 || tile_override() == TILE_64) pb.tile = tile_override();  // tuning (MISTRA_TILE)
    const int64_t te = tile_extent(pb.tile);
    const int64_t slots = int64_t(host_only ? 2 : batch_persistent_occupancy(pb.elem, pb.tile)) * cus;

    // Queue order: longest remaining chain first.  A contraction's rounds run one after the other, so the
    // time from a round's start to its contraction's end is at least the sum of its remaining rounds'
    // durations; ordering items by that sum (descending) starts the long chains early and fills the gaps
    // with short contractions, where round-major order would hold every later round of a large
    // contraction behind all contractions' earlier rounds.  A round's duration is modelled as its
    // slowest product: one tile's K loop plus its C targets' epilogue, times the waves of tiles the
    // resident workgroups need for it.  The sum falls strictly from one round of a contraction to the
    // next, so every item still follows the items it waits for.
    {
        std::map<std::pair<int, int>, double> round_len;  // (entry, round) -> modelled duration
        for (const HostItem& it : items)
        {
            const int64_t tiles = ((it.ms + te - 1) / te) * ((it.ns + te - 1) / te);
            const double waves = std::max(1.0, double(tiles) / double(slots));
            const double d = double(it.kpad + 32 * it.prod.nC + 64) * waves;
            double& r = round_len[{it.entry, it.round}];
            r = std::max(r, d);
        }
        std::map<std::pair<int, int>, double> remaining;  // sum of the round's and later rounds' durations
        double acc = 0.0;
        int cur = -1;
        for (auto it = round_len.rbegin(); it != round_len.rend(); ++it)
        {
            if (it->first.first != cur)
            {
                cur = it->first.first;
                acc = 0.0;
            }
            acc += it->second;
            remaining[it->first] = acc;
        }
        std::stable_sort(items.begin(), items.end(), [&](const HostItem& a, const HostItem& b) {
            const double ra = remaining[{a.entry, a.round}], rb = remaining[{b.entry, b.round}];
            if (ra != rb) return ra > rb;
            if (a.entry != b.entry) return a.entry < b.entry;
            return a.round < b.round;
        });
    }

    // sizes: packed panels, and the table block [packs][items][products][counters][iota][row vectors]
    const int nitems = int(items.size());
    std::vector<int64_t> panel_off(2 * size_t(nitems));
    size_t work_bytes = 0;
    int64_t max_kpad = 0;
    std::map<int64_t, int64_t> rows_of_ld;  // leading dimension -> most rows any panel with it has
    for (int i = 0; i < nitems; i++)
    {
        const HostItem& it = items[size_t(i)];
        panel_off[2 * size_t(i)] = int64_t(work_bytes);
        work_bytes += align256(size_t(it.ms * it.kpad) * size_t(pb.elem));
        panel_off[2 * size_t(i) + 1] = int64_t(work_bytes);
        work_bytes += align256(size_t(it.ns * it.kpad) * size_t(pb.elem));
        max_kpad = std::max(max_kpad, it.kpad);
        int64_t& r = rows_of_ld[it.kpad];
        r = std::max(r, std::max(it.ms, it.ns));
    }
    const int npacks = 2 * nitems;
    const size_t o_packs = 0, o_items = align256(size_t(npacks) * sizeof(PackDesc)) + o_packs;
    const int nsegs = npacks + nitems;
    const size_t o_segs = o_items + align256(size_t(nitems) * sizeof(TileItem));
    const size_t o_prods = o_segs + align256(size_t(nsegs) * sizeof(Segment));
    const size_t o_cnt = o_prods + align256(size_t(nitems) * sizeof(ProdDesc));
    const size_t ncnt = size_t(kPersistHeader + nitems + nentries);
    const size_t o_iota = o_cnt + align256(ncnt * sizeof(unsigned));
    size_t o_rows = o_iota + align256(size_t(max_kpad) * 8);
    std::map<int64_t, size_t> rowvec_off;
    for (const auto& kv : rows_of_ld)
    {
        rowvec_off[kv.first] = o_rows;
        o_rows += align256(size_t(kv.second) * 8);
    }
    const size_t table_bytes = o_rows;

    hipError_t err = hipSuccess;
    if (!host_only)
    {
        err = hipMalloc(&pb.tables, table_bytes);
        if (err == hipSuccess) err = hipMalloc(&pb.work, work_bytes ? work_bytes : 256);
        if (err != hipSuccess)
        {
            (void)hipGetLastError();
            return Status::error(MISTRA_ERR_NOMEM, std::string("persistent batch tables / panels: ") + hipGetErrorString(err));
        }
    }
    // (host_only: placeholder addresses, never dereferenced)
    char* T0 = host_only ? reinterpret_cast<char*>(uintptr_t(1) << 40) : static_cast<char*>(pb.tables);
    char* W0 = host_only ? reinterpret_cast<char*>(uintptr_t(1) << 41) : static_cast<char*>(pb.work);
    std::vector<unsigned char> host(table_bytes, 0);  // counters start at zero
    PackDesc* hp = reinterpret_cast<PackDesc*>(host.data() + o_packs);
    TileItem* hi = reinterpret_cast<TileItem*>(host.data() + o_items);
    Segment* hs = reinterpret_cast<Segment*>(host.data() + o_segs);
    ProdDesc* hd = reinterpret_cast<ProdDesc*>(host.data() + o_prods);
    int64_t* hiota = reinterpret_cast<int64_t*>(host.data() + o_iota);
    for (int64_t k = 0; k < max_kpad; k++) hiota[k] = k;
    for (const auto& kv : rows_of_ld)
    {
        int64_t* v = reinterpret_cast<int64_t*>(host.data() + rowvec_off[kv.first]);
        for (int64_t r = 0; r < kv.second; r++) v[r] = r * kv.first;
    }
    const int64_t* d_iota = reinterpret_cast<const int64_t*>(T0 + o_iota);
    std::vector<int64_t> entry_tiles(size_t(nentries), 0), entry_round_start(size_t(nentries), 0);
    std::vector<int> entry_round(size_t(nentries), -1);
    for (int i = 0; i < nitems; i++)
    {
        const HostItem& it = items[size_t(i)];
        const Plan& p = *it.plan;
        void* pa = W0 + panel_off[2 * size_t(i)];
        void* pbp = W0 + panel_off[2 * size_t(i) + 1];
        // panels: A (rows M through a_m, columns K through a_k), B (rows N through b_n, columns K through b_k)
        for (int w = 0; w < 2; w++)
        {
            PackDesc& d = hp[2 * i + w];
            const bool isA = w == 0;
            d.X = isA ? it.call.A : it.call.B;
            d.out = isA ? pa : pbp;
            d.rs = isA ? p.a_m : p.b_n;
            d.cs = isA ? p.a_k : p.b_k;
            d.rows = isA ? it.ms : it.ns;
            d.cols = it.ks;
            d.ld = it.kpad;
            d.nterms = isA ? it.prod.nA : it.prod.nB;
            for (int t = 0; t < d.nterms; t++)
            {
                const SubRef& r = isA ? it.prod.a[t] : it.prod.b[t];
                d.roff[t] = isA ? r.roff : r.coff;
                d.coff[t] = isA ? r.coff : r.roff;
                d.coef[t] = r.coeff;
            }
            d.rows_fast = isA ? int(p.a_mfast) : int(p.b_nfast);
            d.item = i;
        }
        const int64_t need_pack = pack_jobs_of(it.ms, it.kpad) + pack_jobs_of(it.ns, it.kpad);
        // the product table of the tile jobs: one packed term per operand, the plan's C targets
        ProdDesc& pd = hd[i];
        std::memset(&pd, 0, sizeof(pd));
        pd.nA = pd.nB = 1;
        pd.a[0] = SubRef{0, 0, 0, 1.0};
        pd.b[0] = SubRef{0, 0, 0, 1.0};
        pd.nC = it.prod.nC;
        pd.first_mask = it.prod.first_mask;
        for (int t = 0; t < pd.nC; t++) pd.c[t] = SubRef{0, it.prod.c[t].roff, it.prod.c[t].coff, it.prod.c[t].coeff};
        TileItem& ti = hi[i];
        GemmArgs& g = ti.g;
        std::memset(&g, 0, sizeof(g));
        g.A = pa;
        g.B = pbp;
        g.C = it.call.C;
        const int64_t* rowvec = reinterpret_cast<const int64_t*>(T0 + rowvec_off[it.kpad]);
        g.a_m = rowvec;
        g.b_n = rowvec;
        g.a_k = g.b_k = g.a_krun = g.b_krun = d_iota;
        g.c_m = p.c_m;
        g.c_n = p.c_n;
        g.ms = it.ms;
        g.ns = it.ns;
        g.ks = it.kpad;
        g.prods = reinterpret_cast<const ProdDesc*>(T0 + o_prods) + i;
        g.nprod = 1;
        g.tiles_m = int((it.ms + te - 1) / te);
        g.tiles_n = int((it.ns + te - 1) / te);
        g.alpha = it.alpha;
        g.beta = it.beta;
        g.nbatch = 1;
        g.zero = p.zero;
        g.ka_stride = g.kb_stride = 1;
        g.c_scatter = 1;
        ti.entry = it.entry;
        ti.round = it.round;
        ti.need_pack = need_pack;
        // tiles of the contraction's earlier rounds (items arrive in round order)
        const size_t en = size_t(it.entry);
        if (entry_round[en] != it.round)
        {
            entry_round[en] = it.round;
            entry_round_start[en] = entry_tiles[en];
        }
        ti.need_entry = entry_round_start[en];
        entry_tiles[en] += int64_t(g.tiles_m) * g.tiles_n;
    }
    // Queue positions.  The products' tiles go in the order above; each product's two panels are queued
    // before its tiles, and early enough that about one grid of tile jobs lies between a panel and the
    // first tile that reads it: by the time that tile is taken, its panels are long packed, and from the
    // start the workgroups that pack (memory-bound) run beside workgroups that multiply (MFMA-bound).
    int nseg = 0;
    int64_t pos = 0;
    {
        const int64_t lead = slots;  // tile jobs
        int next = 0;                // next product whose panels are queued
        int64_t covered = 0, queued = 0;
        auto tiles_of = [&](int i) { return int64_t(hi[i].g.tiles_m) * hi[i].g.tiles_n; };
        for (int i = 0; i < nitems; i++)
        {
            while (next < nitems && (next <= i || covered - queued < lead))
            {
                for (int w = 0; w < 2; w++)
                {
                    PackDesc& d = hp[2 * next + w];
                    d.first_job = pos;
                    hs[nseg++] = Segment{pos, 0, 2 * next + w};
                    pos += pack_jobs_of(d.rows, d.ld);
                }
                covered += tiles_of(next);
                next++;
            }
            hi[i].first_job = pos;
            hs[nseg++] = Segment{pos, 1, i};
            pos += tiles_of(i);
            queued += tiles_of(i);
        }
    }
    // a tile's C writes are fenced when a later round of its contraction reads them
    {
        std::vector<int> last_round(size_t(nentries), -1);
        for (int i = 0; i < nitems; i++) last_round[size_t(hi[i].entry)] = std::max(last_round[size_t(hi[i].entry)], hi[i].round);
        for (int i = 0; i < nitems; i++)
        {
            const int last = last_round[size_t(hi[i].entry)];
            hi[i].release = hi[i].round < last ? 1 : 0;
            // every round of a contraction with several rounds reads and writes its C coherently
            hi[i].g.c_coherent = last > 0 ? 1 : 0;
        }
    }
    if (pos + (int64_t(1) << 20) >= (int64_t(1) << 32))
        return Status::error(MISTRA_ERR_UNSUPPORTED, "persistent batch: more than 2^32 jobs");
    for (int64_t n : entry_tiles)
        if (n >= int64_t(1) << 31) return Status::error(MISTRA_ERR_UNSUPPORTED, "persistent batch: 2^31 tiles in one contraction");
    // every launch's queue is replayed on the host before it is uploaded (a few microseconds per thousand
    // segments; a queue that fails it could leave a wait unmet)
    {
        Status qs = check_queue(hs, nseg, hp, npacks, hi, nitems, nentries, pos);
        if (!qs.ok() || host_only) return qs;
    }
    err = hipMemcpy(pb.tables, host.data(), table_bytes, hipMemcpyHostToDevice);
    if (err != hipSuccess) return Status::error(MISTRA_ERR_HIP, std::string("persistent batch upload: ") + hipGetErrorString(err));
    PersistArgs& a = pb.args;
    a.packs = reinterpret_cast<const PackDesc*>(T0 + o_packs);
    a.npacks = npacks;
    a.items = reinterpret_cast<const TileItem*>(T0 + o_items);
    a.nitems = nitems;
    a.segs = reinterpret_cast<const Segment*>(T0 + o_segs);
    a.nsegs = nseg;
    a.nentries = nentries;
    a.njobs = pos;
    a.counters = reinterpret_cast<unsigned*>(T0 + o_cnt);
    a.spin_limit = kDefaultSpin;
    a.dep_bias = 0;
    pb.grid = int(std::max<int64_t>(1, std::min<int64_t>(slots, a.njobs)));
    return Status();
}

// Timed-out waits of the batch's launches so far (host read after the stream drains); clears them.
Status persistent_timeouts(const PersistBatch& pb, hipStream_t stream)
{
    unsigned t = 0;
    hipError_t e = hipStreamSynchronize(stream);
    if (e == hipSuccess) e = hipMemcpy(&t, pb.args.counters + 2, sizeof(t), hipMemcpyDeviceToHost);
    if (e == hipSuccess && t)
    {
        const unsigned zero = 0;
        e = hipMemcpy(pb.args.counters + 2, &zero, sizeof(zero), hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) return Status::error(MISTRA_ERR_HIP, hipGetErrorString(e));
    if (t)
        return Status::error(MISTRA_ERR_HIP, "persistent batch: " + std::to_string(t) +
                                                 " dependency waits gave up (results are not valid)");
    return Status();
}

Status batch_persistent(int count, const mistra_contraction* entries, unsigned flags, hipStream_t stream)
{
    const bool graph = (flags & MISTRA_FLAG_GRAPH) != 0;
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::ostringstream k;
    // per stream: the cached queue counters and panels serve one launch at a time, and calls on one stream
    // are ordered, so calls of the same batch on different streams get their own
    k << dev << "#p" << count << '@' << reinterpret_cast<uintptr_t>(stream);
    for (int i = 0; i < count; i++)
    {
        const mistra_contraction& x = entries[i];
        if (!x.A || !x.B || !x.C) return Status::error(MISTRA_ERR_ARG, "null tensor in batch entry");
        k << '{' << x.level << ',' << x.impl;
        append_tensor(k, x.A, x.idx_A);
        append_tensor(k, x.B, x.idx_B);
        append_tensor(k, x.C, x.idx_C);
        k << '}';
    }
    const std::string key = k.str();
    std::shared_ptr<PersistBatch> pb;
    {
        std::lock_guard<std::mutex> lk(g_persist_mutex);
        for (auto it = g_persist.begin(); it != g_persist.end(); ++it)
            if ((*it)->key == key)
            {
                pb = *it;
                g_persist.splice(g_persist.begin(), g_persist, it);
                break;
            }
    }
    if (!pb)
    {
        // the panels of earlier batches go first if this one does not fit beside them
        auto n = std::make_shared<PersistBatch>();
        n->key = key;
        Status st = prepare_persistent(count, entries, stream, *n);
        if (st.code == MISTRA_ERR_NOMEM)
        {
            {
                std::lock_guard<std::mutex> lk(g_persist_mutex);
                g_persist.clear();
            }
            n = std::make_shared<PersistBatch>();
            n->key = key;
            st = prepare_persistent(count, entries, stream, *n);
        }
        if (!st.ok()) return st;
        pb = n;
        std::lock_guard<std::mutex> lk(g_persist_mutex);
        g_persist.push_front(pb);
        while (g_persist.size() > kMaxPersist) g_persist.pop_back();
    }
    for (int i : pb->singles)
    {
        const mistra_contraction& x = entries[i];
        Status st = tensor_mult(x.A, x.idx_A, x.B, x.idx_B, x.C, x.idx_C, x.level, x.impl, MISTRA_PREC_NATIVE, 0,
                                stream, false);
        if (!st.ok()) return st;
    }
    if (pb->args.njobs > 0)
    {
        PersistArgs a = pb->args;
        const long long spin = g_dbg_batch_spin.load();
        a.spin_limit = spin > 0 ? spin : kDefaultSpin;
        a.dep_bias = g_dbg_batch_bias.load();
        hipError_t e = hipSuccess;
        if (graph && a.spin_limit == kDefaultSpin && a.dep_bias == 0)
        {
            if (!pb->exec)
            {
                e = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal);
                if (e == hipSuccess)
                {
                    const hipError_t el = launch_batch_persistent(pb->elem, pb->tile, a, pb->grid, stream);
                    hipGraph_t gph = nullptr;
                    e = hipStreamEndCapture(stream, &gph);
                    if (e == hipSuccess) e = el;
                    if (e == hipSuccess) pb->graph = gph;
                    if (e == hipSuccess) e = hipGraphInstantiate(&pb->exec, gph, nullptr, nullptr, 0);
                }
                if (e != hipSuccess)
                    return Status::error(MISTRA_ERR_HIP, std::string("persistent batch capture: ") + hipGetErrorString(e));
            }
            timing_begin(stream);
            e = hipGraphLaunch(pb->exec, stream);
            timing_end(stream, pb->flops);
        }
        else
        {
            timing_begin(stream);
            e = launch_batch_persistent(pb->elem, pb->tile, a, pb->grid, stream);
            timing_end(stream, pb->flops);
        }
        if (e != hipSuccess) return Status::error(MISTRA_ERR_HIP, std::string("persistent batch launch: ") + hipGetErrorString(e));
    }
    for (int i = 0; i < count; i++)
    {
        tblis::tblis_tensor* C = entries[i].C;
        set_scalar(C, 1.0);
        C->conj = 0;
    }
    if (flags & MISTRA_FLAG_SYNC)
    {
        if (pb->args.njobs > 0) return persistent_timeouts(*pb, stream);
        hipError_t e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return Status::error(MISTRA_ERR_HIP, hipGetErrorString(e));
    }
    return Status();
}

}  // namespace

Status batch_mult(int count, const mistra_contraction* entries, unsigned flags, hipStream_t stream)
{
    if (count < 0 || (count > 0 && !entries)) return Status::error(MISTRA_ERR_ARG, "bad batch");
    if (!(flags & MISTRA_FLAG_BATCH_GROUPED))
    {
        if ((flags & MISTRA_FLAG_GRAPH) && !stream)
            return Status::error(MISTRA_ERR_ARG, "graph replay needs a non-null stream");
        return batch_persistent(count, entries, flags, stream);
    }
    return batch_grouped(count, entries, flags, stream);
}

namespace {

Status batch_grouped(int count, const mistra_contraction* entries, unsigned flags, hipStream_t stream)
{
    if (count < 0 || (count > 0 && !entries)) return Status::error(MISTRA_ERR_ARG, "bad batch");
    const bool graph = (flags & MISTRA_FLAG_GRAPH) != 0;
    if (graph && !stream) return Status::error(MISTRA_ERR_ARG, "graph replay needs a non-null stream");

    std::string key;
    if (graph)
    {
        int dev = 0;
        (void)hipGetDevice(&dev);
        std::ostringstream k;
        k << dev << '#' << count;
        for (int i = 0; i < count; i++)
        {
            const mistra_contraction& x = entries[i];
            if (!x.A || !x.B || !x.C) return Status::error(MISTRA_ERR_ARG, "null tensor in batch entry");
            k << '{' << x.level << ',' << x.impl;
            append_tensor(k, x.A, x.idx_A);
            append_tensor(k, x.B, x.idx_B);
            append_tensor(k, x.C, x.idx_C);
            k << '}';
        }
        key = k.str();
        std::shared_ptr<BatchGraph> hit;
        {
            std::lock_guard<std::mutex> lk(g_graph_mutex);
            for (auto it = g_graphs.begin(); it != g_graphs.end(); ++it)
                if ((*it)->key == key)
                {
                    hit = *it;
                    g_graphs.splice(g_graphs.begin(), g_graphs, it);
                    break;
                }
        }
        if (hit)
        {
            // entries outside the graph run individually first (they are few: host data, general classes)
            std::vector<std::shared_ptr<Plan>> plans;
            std::vector<int> singles;
            std::vector<GemmArgs> jobs;
            std::vector<int64_t> first;
            std::vector<Group> groups;
            Status st = prepare(count, entries, plans, singles, jobs, first, groups, stream);
            if (!st.ok()) return st;
            for (int i : singles)
            {
                const mistra_contraction& x = entries[i];
                st = tensor_mult(x.A, x.idx_A, x.B, x.idx_B, x.C, x.idx_C, x.level, x.impl, MISTRA_PREC_NATIVE, 0,
                                 stream, false);
                if (!st.ok()) return st;
            }
            hipError_t e = hipGraphLaunch(hit->exec, stream);
            if (e != hipSuccess) return Status::error(MISTRA_ERR_HIP, std::string("graph launch: ") + hipGetErrorString(e));
            for (int i = 0; i < count; i++)
            {
                // as after any call, the scalars of C read 1
                tblis::tblis_tensor* C = entries[i].C;
                if (C->type == tblis::TYPE_DOUBLE)
                {
                    const double one = 1.0;
                    std::memcpy(C->scalar, &one, sizeof(one));
                }
                else
                {
                    const float one = 1.0f;
                    std::memcpy(C->scalar, &one, sizeof(one));
                }
            }
            if (flags & MISTRA_FLAG_SYNC)
            {
                e = hipStreamSynchronize(stream);
                if (e != hipSuccess) return Status::error(MISTRA_ERR_HIP, hipGetErrorString(e));
            }
            return Status();
        }
    }

    std::vector<std::shared_ptr<Plan>> plans;
    std::vector<int> singles;
    std::vector<GemmArgs> jobs;
    std::vector<int64_t> first;
    std::vector<Group> groups;
    Status st = prepare(count, entries, plans, singles, jobs, first, groups, stream);
    if (!st.ok()) return st;
    for (int i : singles)
    {
        const mistra_contraction& x = entries[i];
        st = tensor_mult(x.A, x.idx_A, x.B, x.idx_B, x.C, x.idx_C, x.level, x.impl, MISTRA_PREC_NATIVE, 0, stream,
                         false);
        if (!st.ok()) return st;
    }
    if (!groups.empty())
    {
        if (graph)
        {
            auto bg = std::make_shared<BatchGraph>();
            bg->key = key;
            bg->plans = plans;
            size_t first_base = 0;
            st = upload_tables(jobs, first, bg->tables, first_base, stream, false);
            if (!st.ok()) return st;
            hipError_t e = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal);
            if (e != hipSuccess) return Status::error(MISTRA_ERR_HIP, std::string("capture: ") + hipGetErrorString(e));
            st = launch_groups(groups, bg->tables, first_base, stream, false);
            hipGraph_t gph = nullptr;
            e = hipStreamEndCapture(stream, &gph);
            if (!st.ok()) return st;
            if (e != hipSuccess) return Status::error(MISTRA_ERR_HIP, std::string("capture end: ") + hipGetErrorString(e));
            bg->graph = gph;
            e = hipGraphInstantiate(&bg->exec, gph, nullptr, nullptr, 0);
            if (e != hipSuccess) return Status::error(MISTRA_ERR_HIP, std::string("instantiate: ") + hipGetErrorString(e));
            e = hipGraphLaunch(bg->exec, stream);
            if (e != hipSuccess) return Status::error(MISTRA_ERR_HIP, std::string("graph launch: ") + hipGetErrorString(e));
            std::lock_guard<std::mutex> lk(g_graph_mutex);
            g_graphs.push_front(bg);
            while (g_graphs.size() > kMaxGraphs) g_graphs.pop_back();
        }
        else
        {
            void* tables = nullptr;
            size_t first_base = 0;
            st = upload_tables(jobs, first, tables, first_base, stream, true);
            if (!st.ok()) return st;
            st = launch_groups(groups, tables, first_base, stream, true);
            (void)hipFreeAsync(tables, stream);
            if (!st.ok()) return st;
        }
    }
    for (int i = 0; i < count; i++)
    {
        tblis::tblis_tensor* C = entries[i].C;
        if (C->type == tblis::TYPE_DOUBLE)
        {
            const double one = 1.0;
            std::memcpy(C->scalar, &one, sizeof(one));
        }
        else
        {
            const float one = 1.0f;
            std::memcpy(C->scalar, &one, sizeof(one));
        }
        C->conj = 0;
    }
    if (flags & MISTRA_FLAG_SYNC)
    {
        hipError_t e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return Status::error(MISTRA_ERR_HIP, hipGetErrorString(e));
    }
    return Status();
}

}  // namespace

Status batch_check_host(int count, const mistra_contraction* entries)
{
    PersistBatch pb;
    return prepare_persistent(count, entries, nullptr, pb, true);
}

void release_batch_graphs()
{
    {
        std::lock_guard<std::mutex> lk(g_graph_mutex);
        g_graphs.clear();
    }
    std::lock_guard<std::mutex> lk(g_persist_mutex);
    g_persist.clear();
}

}  // namespace mistra
