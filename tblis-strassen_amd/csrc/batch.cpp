// Batched contractions (mistra_batch_mult): many independent contractions with few launches.
//
// Every device-resident ABC contraction contributes its launches (Strassen stages, then fringe passes)
// as jobs with a round number; the jobs of one round and one kernel variant become a single grouped
// launch (leaf_gemm_grouped: a 1-D grid whose workgroups look their job up in a device table).  Rounds
// run in stream order, which keeps every contraction's stage order.  With MISTRA_FLAG_GRAPH the
// launch sequence is captured once into a HIP graph (kernel nodes only: the tables are uploaded before
// capture and owned by the cache entry) and replayed while the entries stay the same.
#include <cstring>
#include <list>
#include <map>
#include <mutex>
#include <sstream>
#include <tuple>

#include "driver.hpp"

namespace mistra {

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
    hipError_t e = async ? hipMallocAsync(&buf, total ? total : 1, stream) : hipMalloc(&buf, total ? total : 1);
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

}  // namespace

Status batch_mult(int count, const mistra_contraction* entries, unsigned flags, hipStream_t stream)
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

void release_batch_graphs()
{
    std::lock_guard<std::mutex> lk(g_graph_mutex);
    g_graphs.clear();
}

}  // namespace mistra
