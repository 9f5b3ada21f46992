// Host-plan introspection for tests (include/tblis_strassen_debug.h); never touches the device.
#include <cstring>

#include "driver.hpp"
#include "tblis_strassen_debug.h"

namespace tblis {
extern "C" {

int mistra_debug_plan(const tblis_tensor* A, const label_type* idx_A, const tblis_tensor* B, const label_type* idx_B,
                      const tblis_tensor* C, const label_type* idx_C, int level, unsigned flags, long long* dims,
                      long long* scat, long long cap, double* prods, int cap_prods)
{
    mistra::Plan p;
    mistra::Status st = mistra::build_plan_host(A, idx_A, B, idx_B, C, idx_C, level, mistra::IMPL_ABC,
                                                -1, flags, p);
    if (!st.ok()) return st.code;
    std::vector<int> order;
    mistra::make_stages(p, order);
    const long long d[16] = {p.m, p.n, p.k, p.nbatch, p.general ? 0 : 1, p.transposed, p.ms, p.ns, p.ks,
                             p.nprod, (long long)p.fringe.size(), p.a_mfast, p.b_nfast, p.staged_depth,
                             (long long)p.h_stage_begin.size() - 1, p.outer_staged};
    std::memcpy(dims, d, sizeof(d));
    const std::vector<int64_t>* v[6] = {&p.h_a_m, &p.h_a_k, &p.h_b_k, &p.h_b_n, &p.h_c_m, &p.h_c_n};
    for (int i = 0; i < 6; i++)
        for (long long j = 0; j < cap; j++)
            scat[i * cap + j] = j < (long long)v[i]->size() ? (*v[i])[size_t(j)] : -1;
    const int np = int(p.h_prods.size());
    for (int q = 0; q < np && q < cap_prods; q++)
    {
        const mistra::ProdDesc& pd = p.h_prods[size_t(q)];
        double* o = prods + q * 40;
        o[0] = pd.nA;
        o[1] = pd.nB;
        o[2] = pd.nC;
        o[3] = pd.first_mask;
        const mistra::SubRef* refs[3] = {pd.a, pd.b, pd.c};
        const int cnt[3] = {pd.nA, pd.nB, pd.nC};
        for (int s = 0; s < 3; s++)
            for (int t = 0; t < 4; t++)
            {
                double* r = o + 4 + s * 12 + t * 3;
                if (t < cnt[s])
                {
                    r[0] = double(refs[s][t].roff);
                    r[1] = double(refs[s][t].coff);
                    r[2] = refs[s][t].coeff;
                }
                else
                    r[0] = r[1] = -1, r[2] = 0;
            }
    }
    return np;
}

int mistra_debug_run_strides(const long long* scat, long long n, int run, long long* out)
{
    std::vector<int64_t> v(scat, scat + n);
    std::vector<int64_t> r = mistra::run_strides(v, run);
    for (long long i = 0; i < n; i++) out[i] = r[size_t(i)];
    return 0;
}

int mistra_debug_set_combine_form(int per_sub_block)
{
    return mistra::g_combine_per_sub.exchange(per_sub_block ? 1 : 0);
}

int mistra_debug_batch_check(int count, const mistra_contraction* entries, char* msg, int msg_len)
{
    const mistra::Status st = mistra::batch_check_host(count, entries);
    if (msg && msg_len > 0)
    {
        std::strncpy(msg, st.msg.c_str(), size_t(msg_len) - 1);
        msg[msg_len - 1] = 0;
    }
    return st.code;
}

int mistra_debug_set_batch_wait(long long spin_limit, unsigned dep_bias)
{
    mistra::g_dbg_batch_spin.store(spin_limit > 0 ? spin_limit : 0);
    mistra::g_dbg_batch_bias.store(dep_bias);
    return 0;
}

int mistra_debug_set_lowp(int bf16_slices, int int8_slices, int int8_fold_ksteps)
{
    if (bf16_slices != 0 && (bf16_slices < 2 || bf16_slices > 7)) return MISTRA_ERR_ARG;
    if (int8_slices != 0 && (int8_slices < 2 || int8_slices > 6)) return MISTRA_ERR_ARG;
    if (int8_fold_ksteps < 0 || int8_fold_ksteps > 1024) return MISTRA_ERR_ARG;
    mistra::g_dbg_ozaki_slices.store(bf16_slices);
    mistra::g_dbg_i8_slices.store(int8_slices);
    mistra::g_dbg_i8_fold.store(int8_fold_ksteps);
    return 0;
}

int mistra_debug_tune_candidates(const tblis_tensor* A, const label_type* idx_A, const tblis_tensor* B,
                                 const label_type* idx_B, const tblis_tensor* C, const label_type* idx_C, int* level,
                                 int* impl, int* tile, double* model_ms, int cap)
{
    std::vector<mistra::TuneCand> cands;
    bool general = false;
    const mistra::Status st = mistra::tune_candidates(A, idx_A, B, idx_B, C, idx_C, cands, general);
    if (!st.ok()) return st.code;
    const int n = int(cands.size());
    for (int i = 0; i < n && i < cap; i++)
    {
        level[i] = cands[size_t(i)].level;
        impl[i] = cands[size_t(i)].impl;
        tile[i] = cands[size_t(i)].tile;
        model_ms[i] = cands[size_t(i)].ms;
    }
    return n;
}

int mistra_debug_set_lowp_fusion(int fused)
{
    mistra::g_dbg_unfused_lowp.store(fused ? 0 : 1);
    return 0;
}

int mistra_debug_set_tile(int tile)
{
    if (tile < -1 || tile >= mistra::kNumTiles) return MISTRA_ERR_ARG;
    mistra::g_dbg_tile.store(tile);
    return 0;
}

}  // extern "C"
}  // namespace tblis
