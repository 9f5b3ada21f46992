/*
 * oracle.c -- CPU restatement of flame/tblis-strassen's Strassen tensor contraction, FP64.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker (or, for cpu_baseline, as the timed CPU port).  The
 * product (tblis-strassen_amd/) never links or calls it.
 *
 * Functions and the reference code they restate:
 *   oracle_contract_ref   brute-force contraction over every label assignment; the classical triple
 *                         loop of contract_ref (src/internal/3t/mult.cxx:101-165) generalised to all
 *                         seven index classes (A-only/B-only summed, C-only broadcast, batch)
 *   ix_diagonal           diagonal()        src/util/tensor.hpp:305-342
 *   ix_group              index grouping    src/stra_iface/3t/stra_mult.cxx:35-180
 *   ix_fold               fold()            src/util/tensor.hpp:270-303
 *   ix_sort2              sort_by_stride()  src/util/tensor.hpp:37-78, used as in
 *                                           src/stra_internal/3t/stra_mult.cxx:336-356
 *   ix_scatter            fill_scatter()    src/matrix/stra_tensor_view.hpp:427-455
 *   strassen tables       src/stra_internal/3t/strassen:1-14 (U,V,W), products in the order of
 *                         src/stra_internal/3t/stra_mult.cxx:426-591; 2-level = Kronecker square in
 *                         the order of src/stra_internal/3t/stra_mult_2level.cxx:341-1458
 *   oracle_tensor_mult_blocked  the ABC impl through the reference's blocked loop nest (see below); the
 *                         timed CPU baseline
 *   oracle_tensor_mult    internal::stra_mult dispatch (src/stra_internal/3t/stra_mult.cxx:1701-1846)
 *                         with the three impls:
 *                           ABC    sums fused into packing, per-KC-panel C_i += coeff*AB
 *                                  (stra_contract_blis :308-605, src/stra_nodes/stra_partm.hpp:98,
 *                                   src/stra_nodes/stra_gemm_ukr.hpp:626-930), KC = 256
 *                           AB     M = alpha*(sum A)(sum B) into a temporary, then C_i += coeff*M
 *                                  (straprim_ab2, src/stra_internal/3t/straprim_common.hpp:160-204)
 *                           Naive  explicit S, T, M = alpha*S*T, then C_i += coeff*M
 *                                  (straprim_naive2, src/stra_internal/3t/straprim_common.hpp:70-157)
 *                         The remainder of the 2^L split is dropped when ref_fringe != 0, as
 *                         src/matrix/stra_tensor_view.hpp:246-247 does; otherwise it is computed by
 *                         classical passes.  alpha/beta follow BLAS semantics (beta applied once).
 *
 * Arithmetic: FP64, products then sums (the build uses -ffp-contract=off, matching the mul-then-add
 * of the reference's AVX kernels).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXD 16

typedef struct
{
    int n;
    char idx[MAXD];
    long len[MAXD];
    long sA[MAXD], sB[MAXD], sC[MAXD];
} group_t;

/* ---------------------------------------------------------------------------------------------- */
/* brute force                                                                                     */
/* ---------------------------------------------------------------------------------------------- */

static int find_label(const char* labels, int n, char l)
{
    for (int i = 0; i < n; i++)
        if (labels[i] == l) return i;
    return -1;
}

/* offset of a tensor element for a full label assignment (repeated labels add their strides) */
static long offset_of(int nd, const long* stride, const char* idx, const char* labels, int nl, const long* val)
{
    long off = 0;
    for (int d = 0; d < nd; d++) off += stride[d] * val[find_label(labels, nl, idx[d])];
    return off;
}

int oracle_contract_ref(int ndA, const long* lenA, const long* strideA, const char* idxA, const double* A,
                        int ndB, const long* lenB, const long* strideB, const char* idxB, const double* B,
                        int ndC, const long* lenC, const long* strideC, const char* idxC, double* C,
                        double alpha, double beta)
{
    char labels[3 * MAXD];
    long L[3 * MAXD];
    int nl = 0;
    const int nds[3] = {ndA, ndB, ndC};
    const long* lens[3] = {lenA, lenB, lenC};
    const char* idxs[3] = {idxA, idxB, idxC};
    for (int t = 0; t < 3; t++)
        for (int d = 0; d < nds[t]; d++)
        {
            int f = find_label(labels, nl, idxs[t][d]);
            if (f < 0)
            {
                labels[nl] = idxs[t][d];
                L[nl++] = lens[t][d];
            }
            else if (L[f] != lens[t][d])
                return -1;
        }
    /* C labels (distinct) and the summed labels (not in C) */
    int cl[3 * MAXD], ncl = 0, sl[3 * MAXD], nsl = 0;
    for (int i = 0; i < nl; i++)
    {
        if (find_label(idxC, ndC, labels[i]) >= 0) cl[ncl++] = i;
        else sl[nsl++] = i;
    }
    long val[3 * MAXD];
    memset(val, 0, sizeof(val));
    for (;;)
    {
        /* one C element */
        long offC = offset_of(ndC, strideC, idxC, labels, nl, val);
        double acc = 0.0;
        int empty = 0;
        for (int i = 0; i < nsl; i++)
            if (L[sl[i]] == 0) empty = 1;
        for (int i = 0; i < nsl; i++) val[sl[i]] = 0;
        while (!empty)
        {
            double a = A[offset_of(ndA, strideA, idxA, labels, nl, val)];
            double b = B[offset_of(ndB, strideB, idxB, labels, nl, val)];
            acc += a * b;
            int i = 0;
            for (; i < nsl; i++)
            {
                if (++val[sl[i]] < L[sl[i]]) break;
                val[sl[i]] = 0;
            }
            if (i == nsl) break;
        }
        C[offC] = (beta == 0.0 ? 0.0 : beta * C[offC]) + alpha * acc;
        int i = 0;
        for (; i < ncl; i++)
        {
            if (++val[cl[i]] < L[cl[i]]) break;
            val[cl[i]] = 0;
        }
        if (i == ncl) break;
        for (int j = 0; j < ncl; j++)
            if (L[cl[j]] == 0) return 0;
    }
    return 0;
}

/* ---------------------------------------------------------------------------------------------- */
/* index layer                                                                                     */
/* ---------------------------------------------------------------------------------------------- */

/* sort dims by label, merge repeated labels by summing strides, drop length-1 dims */
static int ix_diagonal(int nd, const long* len, const long* stride, const char* idx, long* lo, long* so, char* io)
{
    int ord[MAXD];
    for (int i = 0; i < nd; i++) ord[i] = i;
    for (int i = 1; i < nd; i++) /* insertion sort: stable */
        for (int j = i; j > 0 && idx[ord[j - 1]] > idx[ord[j]]; j--)
        {
            int t = ord[j];
            ord[j] = ord[j - 1];
            ord[j - 1] = t;
        }
    int n = 0;
    for (int i = 0; i < nd; i++)
    {
        int d = ord[i];
        if (len[d] == 1) continue;
        if (n > 0 && io[n - 1] == idx[d]) so[n - 1] += stride[d];
        else
        {
            lo[n] = len[d];
            so[n] = stride[d];
            io[n] = idx[d];
            n++;
        }
    }
    return n;
}

static void ix_sort2(int n, const long* s1, const long* s2, int* ord)
{
    for (int i = 0; i < n; i++) ord[i] = i;
    for (int i = 1; i < n; i++)
        for (int j = i; j > 0; j--)
        {
            int a = ord[j - 1], b = ord[j];
            int gt = s1[a] > s1[b] || (s1[a] == s1[b] && s2 && s2[a] > s2[b]);
            if (!gt) break;
            ord[j - 1] = b;
            ord[j] = a;
        }
}

static void ix_permute(group_t* g, const int* ord)
{
    group_t o = *g;
    for (int i = 0; i < g->n; i++)
    {
        o.idx[i] = g->idx[ord[i]];
        o.len[i] = g->len[ord[i]];
        o.sA[i] = g->sA[ord[i]];
        o.sB[i] = g->sB[ord[i]];
        o.sC[i] = g->sC[ord[i]];
    }
    *g = o;
}

/* fold: order by the first used stride vector; merge i into the previous kept index when contiguous
   in every used stride vector.  use = bitmask of used vectors (1=A, 2=B, 4=C). */
static void ix_fold(group_t* g, int use)
{
    if (g->n == 0) return;
    const long* first = (use & 1) ? g->sA : (use & 2) ? g->sB : g->sC;
    int ord[MAXD];
    ix_sort2(g->n, first, NULL, ord);
    group_t o;
    memset(&o, 0, sizeof(o));
    for (int i = 0; i < g->n; i++)
    {
        int d = ord[i];
        int contiguous = 0;
        if (i > 0)
        {
            int p = ord[i - 1];
            contiguous = 1;
            if ((use & 1) && g->sA[d] != g->sA[p] * g->len[p]) contiguous = 0;
            if ((use & 2) && g->sB[d] != g->sB[p] * g->len[p]) contiguous = 0;
            if ((use & 4) && g->sC[d] != g->sC[p] * g->len[p]) contiguous = 0;
        }
        if (contiguous) o.len[o.n - 1] *= g->len[d];
        else
        {
            o.idx[o.n] = g->idx[d];
            o.len[o.n] = g->len[d];
            o.sA[o.n] = g->sA[d];
            o.sB[o.n] = g->sB[d];
            o.sC[o.n] = g->sC[d];
            o.n++;
        }
    }
    *g = o;
}

/* position -> offset over a group with the first index fastest */
static long* ix_scatter(const group_t* g, const long* s, long* count)
{
    long n = 1;
    for (int d = 0; d < g->n; d++) n *= g->len[d];
    long* out = (long*)malloc(sizeof(long) * (size_t)(n > 0 ? n : 1));
    for (long p = 0; p < n; p++)
    {
        long rem = p, off = 0;
        for (int d = 0; d < g->n; d++)
        {
            off += (rem % g->len[d]) * s[d];
            rem /= g->len[d];
        }
        out[p] = off;
    }
    *count = n;
    return out;
}

typedef struct
{
    group_t AB, AC, BC, ABC, Ao, Bo, Co;
} groups_t;

static int ix_group(int ndA, const long* lenA, const long* strideA, const char* idxA, int ndB, const long* lenB,
                    const long* strideB, const char* idxB, int ndC, const long* lenC, const long* strideC,
                    const char* idxC, groups_t* G)
{
    long la[MAXD], sa[MAXD], lb[MAXD], sb[MAXD], lc[MAXD], sc[MAXD];
    char ia[MAXD], ib[MAXD], ic[MAXD];
    int na = ix_diagonal(ndA, lenA, strideA, idxA, la, sa, ia);
    int nb = ix_diagonal(ndB, lenB, strideB, idxB, lb, sb, ib);
    int nc = ix_diagonal(ndC, lenC, strideC, idxC, lc, sc, ic);
    memset(G, 0, sizeof(*G));
    /* visit labels in ascending order (the set operations of the reference sort) */
    for (int c = 0; c < 256; c++)
    {
        char l = (char)c;
        int fa = find_label(ia, na, l), fb = find_label(ib, nb, l), fc = find_label(ic, nc, l);
        if (fa < 0 && fb < 0 && fc < 0) continue;
        long L = fa >= 0 ? la[fa] : fb >= 0 ? lb[fb] : lc[fc];
        if ((fa >= 0 && la[fa] != L) || (fb >= 0 && lb[fb] != L) || (fc >= 0 && lc[fc] != L)) return -1;
        group_t* g = fa >= 0 && fb >= 0 && fc >= 0 ? &G->ABC
                     : fa >= 0 && fb >= 0       ? &G->AB
                     : fa >= 0 && fc >= 0       ? &G->AC
                     : fb >= 0 && fc >= 0       ? &G->BC
                     : fa >= 0                  ? &G->Ao
                     : fb >= 0                  ? &G->Bo
                                                : &G->Co;
        g->idx[g->n] = l;
        g->len[g->n] = L;
        g->sA[g->n] = fa >= 0 ? sa[fa] : 0;
        g->sB[g->n] = fb >= 0 ? sb[fb] : 0;
        g->sC[g->n] = fc >= 0 ? sc[fc] : 0;
        g->n++;
    }
    ix_fold(&G->ABC, 7);
    ix_fold(&G->AB, 3);
    ix_fold(&G->AC, 5);
    ix_fold(&G->BC, 6);
    ix_fold(&G->Ao, 1);
    ix_fold(&G->Bo, 2);
    ix_fold(&G->Co, 4);
    return 0;
}

/* ---------------------------------------------------------------------------------------------- */
/* Strassen tables                                                                                 */
/* ---------------------------------------------------------------------------------------------- */

/* U, V, W of src/stra_internal/3t/strassen (rows = sub-blocks 0..3 of A / B / C, cols = products) */
static const int U1[4][7] = {{1, 0, 1, 0, 1, -1, 0}, {0, 0, 0, 0, 1, 0, 1}, {0, 1, 0, 0, 0, 1, 0}, {1, 1, 0, 1, 0, 0, -1}};
static const int V1[4][7] = {{1, 1, 0, -1, 0, 1, 0}, {0, 0, 1, 0, 0, 1, 0}, {0, 0, 0, 1, 0, 0, 1}, {1, 0, -1, 0, 1, 0, 1}};
static const int W1[4][7] = {{1, 0, 0, 1, -1, 0, 1}, {0, 0, 1, 0, 1, 0, 0}, {0, 1, 0, 1, 0, 0, 0}, {1, -1, 1, 0, 0, 1, 0}};

typedef struct
{
    int n;
    int sub[4];
    double coef[4];
} terms_t;

typedef struct
{
    terms_t A, B, C;
} prod_t;

static void terms_from_column(const int T[4][7], int p, terms_t* t)
{
    t->n = 0;
    for (int s = 0; s < 4; s++)
        if (T[s][p])
        {
            t->sub[t->n] = s;
            t->coef[t->n] = T[s][p];
            t->n++;
        }
}

static void kron_terms(const terms_t* o, const terms_t* i, terms_t* out)
{
    out->n = 0;
    for (int a = 0; a < o->n; a++)
        for (int b = 0; b < i->n; b++)
        {
            int s1 = o->sub[a] / 2, t1 = o->sub[a] % 2, s2 = i->sub[b] / 2, t2 = i->sub[b] % 2;
            out->sub[out->n] = 4 * (2 * s1 + s2) + (2 * t1 + t2);
            out->coef[out->n] = o->coef[a] * i->coef[b];
            out->n++;
        }
}

static int make_products(int level, prod_t* P)
{
    prod_t one[7];
    for (int p = 0; p < 7; p++)
    {
        terms_from_column(U1, p, &one[p].A);
        terms_from_column(V1, p, &one[p].B);
        terms_from_column(W1, p, &one[p].C);
    }
    if (level == 0)
    {
        P[0].A.n = P[0].B.n = P[0].C.n = 1;
        P[0].A.sub[0] = P[0].B.sub[0] = P[0].C.sub[0] = 0;
        P[0].A.coef[0] = P[0].B.coef[0] = P[0].C.coef[0] = 1.0;
        return 1;
    }
    if (level == 1)
    {
        memcpy(P, one, sizeof(one));
        return 7;
    }
    for (int po = 0; po < 7; po++)
        for (int pi = 0; pi < 7; pi++)
        {
            kron_terms(&one[po].A, &one[pi].A, &P[7 * po + pi].A);
            kron_terms(&one[po].B, &one[pi].B, &P[7 * po + pi].B);
            kron_terms(&one[po].C, &one[pi].C, &P[7 * po + pi].C);
        }
    return 49;
}

/* Flattened table for tests: per product, for A, B, C: n, then n (subid, coeff) pairs, padded to 4.
   out must hold nprod * 3 * 9 doubles.  Returns nprod. */
int oracle_products(int level, double* out)
{
    prod_t P[49];
    int np = make_products(level, P);
    for (int p = 0; p < np; p++)
    {
        terms_t* ts[3] = {&P[p].A, &P[p].B, &P[p].C};
        for (int o = 0; o < 3; o++)
        {
            double* q = out + (p * 3 + o) * 9;
            q[0] = ts[o]->n;
            for (int i = 0; i < 4; i++)
            {
                q[1 + 2 * i] = i < ts[o]->n ? ts[o]->sub[i] : -1;
                q[2 + 2 * i] = i < ts[o]->n ? ts[o]->coef[i] : 0;
            }
        }
    }
    return np;
}

/* ---------------------------------------------------------------------------------------------- */
/* Strassen contraction                                                                            */
/* ---------------------------------------------------------------------------------------------- */

typedef struct
{
    long m, n, k;
    long *am, *ak, *bk, *bn, *cm, *cn;
    long nbat;
    long *ba, *bb, *bc;
} flat_t;

static void flat_free(flat_t* f)
{
    free(f->am);
    free(f->ak);
    free(f->bk);
    free(f->bn);
    free(f->cm);
    free(f->cn);
    free(f->ba);
    free(f->bb);
    free(f->bc);
}

static int flatten(int ndA, const long* lenA, const long* strideA, const char* idxA, int ndB, const long* lenB,
                   const long* strideB, const char* idxB, int ndC, const long* lenC, const long* strideC,
                   const char* idxC, flat_t* f, groups_t* G)
{
    if (ix_group(ndA, lenA, strideA, idxA, ndB, lenB, strideB, idxB, ndC, lenC, strideC, idxC, G)) return -1;
    int ord[MAXD];
    ix_sort2(G->AC.n, G->AC.sC, G->AC.sA, ord);
    ix_permute(&G->AC, ord);
    ix_sort2(G->BC.n, G->BC.sC, G->BC.sB, ord);
    ix_permute(&G->BC, ord);
    ix_sort2(G->AB.n, G->AB.sA, G->AB.sB, ord);
    ix_permute(&G->AB, ord);
    long c;
    f->am = ix_scatter(&G->AC, G->AC.sA, &f->m);
    f->cm = ix_scatter(&G->AC, G->AC.sC, &c);
    f->ak = ix_scatter(&G->AB, G->AB.sA, &f->k);
    f->bk = ix_scatter(&G->AB, G->AB.sB, &c);
    f->bn = ix_scatter(&G->BC, G->BC.sB, &f->n);
    f->cn = ix_scatter(&G->BC, G->BC.sC, &c);
    f->ba = ix_scatter(&G->ABC, G->ABC.sA, &f->nbat);
    f->bb = ix_scatter(&G->ABC, G->ABC.sB, &c);
    f->bc = ix_scatter(&G->ABC, G->ABC.sC, &c);
    return 0;
}

/* Description of the flattened problem for tests: sizes and the first entries of each scatter. */
int oracle_describe(int ndA, const long* lenA, const long* strideA, const char* idxA, int ndB, const long* lenB,
                    const long* strideB, const char* idxB, int ndC, const long* lenC, const long* strideC,
                    const char* idxC, long* dims /* m,n,k,nbatch,pure */, long* scat /* 6 vectors, cap each */,
                    long cap)
{
    flat_t f;
    groups_t G;
    memset(&f, 0, sizeof(f));
    if (flatten(ndA, lenA, strideA, idxA, ndB, lenB, strideB, idxB, ndC, lenC, strideC, idxC, &f, &G)) return -1;
    dims[0] = f.m;
    dims[1] = f.n;
    dims[2] = f.k;
    dims[3] = f.nbat;
    dims[4] = G.Ao.n == 0 && G.Bo.n == 0 && G.Co.n == 0;
    long* v[6] = {f.am, f.ak, f.bk, f.bn, f.cm, f.cn};
    long cnt[6] = {f.m, f.k, f.k, f.n, f.m, f.n};
    for (int i = 0; i < 6; i++)
        for (long j = 0; j < cap; j++) scat[i * cap + j] = j < cnt[i] ? v[i][j] : -1;
    flat_free(&f);
    return 0;
}

#define KC 256

/* One leaf product on the flattened problem (batch entry z):
   S = sum_a c_a A_a (ms x ks), T = sum_b c_b B_b (ks x ns), then per impl update the C targets. */
static void leaf_product(const flat_t* f, long z, const double* A, const double* B, double* C, const prod_t* P,
                         long d, long ms, long ns, long ks, int impl, double alpha, const int* first, double beta,
                         double* S, double* T, double* M)
{
    const double* Az = A + f->ba[z];
    const double* Bz = B + f->bb[z];
    double* Cz = C + f->bc[z];
    /* explicit sums (the fused packing computes the same values, panel by panel) */
#pragma omp parallel for schedule(static)
    for (long kk = 0; kk < ks; kk++)
        for (long i = 0; i < ms; i++)
        {
            double v = 0.0;
            for (int t = 0; t < P->A.n; t++)
            {
                long r = (P->A.sub[t] / d) * ms + i, c = (P->A.sub[t] % d) * ks + kk;
                v += P->A.coef[t] * Az[f->am[r] + f->ak[c]];
            }
            S[kk * ms + i] = v;
        }
#pragma omp parallel for schedule(static)
    for (long j = 0; j < ns; j++)
        for (long kk = 0; kk < ks; kk++)
        {
            double v = 0.0;
            for (int t = 0; t < P->B.n; t++)
            {
                long r = (P->B.sub[t] / d) * ks + kk, c = (P->B.sub[t] % d) * ns + j;
                v += P->B.coef[t] * Bz[f->bk[r] + f->bn[c]];
            }
            T[j * ks + kk] = v;
        }
    /* products: ABC updates C per KC panel; AB / Naive form the whole M first */
    const long npanel = impl == 0 ? (ks + KC - 1) / KC : 1;
    const long kc = impl == 0 ? KC : ks;
    for (long pc = 0; pc < npanel; pc++)
    {
        const long k0 = pc * kc, k1 = k0 + kc < ks ? k0 + kc : ks;
#pragma omp parallel for schedule(static)
        for (long j = 0; j < ns; j++)
        {
            double* mj = M + j * ms;
            for (long i = 0; i < ms; i++) mj[i] = 0.0;
            for (long kk = k0; kk < k1; kk++)
            {
                const double t = T[j * ks + kk];
                const double* sk = S + kk * ms;
                for (long i = 0; i < ms; i++) mj[i] += sk[i] * t;
            }
            for (int t = 0; t < P->C.n; t++)
            {
                const long r0 = (P->C.sub[t] / d) * ms, c0 = (P->C.sub[t] % d) * ns;
                const double cf = alpha * P->C.coef[t];
                const int apply_beta = first[t] && pc == 0;
                double* cj = Cz + f->cn[c0 + j];
                for (long i = 0; i < ms; i++)
                {
                    double* cp = cj + f->cm[r0 + i];
                    const double v = cf * mj[i];
                    if (apply_beta) *cp = (beta == 0.0 ? 0.0 : beta * *cp) + v;
                    else *cp += v;
                }
            }
        }
    }
}

/* classical pass over rows [r0,r0+mm) x cols [c0,c0+nn) x k [k0,k0+kk) of the flattened problem */
static void classical_pass(const flat_t* f, long z, const double* A, const double* B, double* C, long r0, long mm,
                           long c0, long nn, long k0, long kk, double alpha, double beta, int apply_beta)
{
    const double* Az = A + f->ba[z];
    const double* Bz = B + f->bb[z];
    double* Cz = C + f->bc[z];
#pragma omp parallel for schedule(static)
    for (long j = 0; j < nn; j++)
        for (long i = 0; i < mm; i++)
        {
            double acc = 0.0;
            for (long p = 0; p < kk; p++) acc += Az[f->am[r0 + i] + f->ak[k0 + p]] * Bz[f->bk[k0 + p] + f->bn[c0 + j]];
            double* cp = Cz + f->cm[r0 + i] + f->cn[c0 + j];
            if (apply_beta) *cp = (beta == 0.0 ? 0.0 : beta * *cp) + alpha * acc;
            else *cp += alpha * acc;
        }
}

int oracle_tensor_mult(int level, int impl, int ref_fringe, int nthreads, int ndA, const long* lenA,
                       const long* strideA, const char* idxA, const double* A, int ndB, const long* lenB,
                       const long* strideB, const char* idxB, const double* B, int ndC, const long* lenC,
                       const long* strideC, const char* idxC, double* C, double alpha, double beta)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    flat_t f;
    groups_t G;
    memset(&f, 0, sizeof(f));
    if (flatten(ndA, lenA, strideA, idxA, ndB, lenB, strideB, idxB, ndC, lenC, strideC, idxC, &f, &G)) return -1;
    if (G.Ao.n || G.Bo.n || G.Co.n)
    {
        flat_free(&f);
        /* general index classes: the brute-force definition */
        return oracle_contract_ref(ndA, lenA, strideA, idxA, A, ndB, lenB, strideB, idxB, B, ndC, lenC, strideC,
                                   idxC, C, alpha, beta);
    }
    const long d = 1L << level, ms = f.m / d, ns = f.n / d, ks = f.k / d;
    const int core = level > 0 && ms > 0 && ns > 0 && ks > 0;
    if (!core)
    {
        if (!(level > 0 && ref_fringe))
            for (long z = 0; z < f.nbat; z++) classical_pass(&f, z, A, B, C, 0, f.m, 0, f.n, 0, f.k, alpha, beta, 1);
        flat_free(&f);
        return 0;
    }
    prod_t P[49];
    int np = make_products(level, P);
    double* S = (double*)malloc(sizeof(double) * (size_t)(ms * ks > 0 ? ms * ks : 1));
    double* T = (double*)malloc(sizeof(double) * (size_t)(ks * ns > 0 ? ks * ns : 1));
    double* M = (double*)malloc(sizeof(double) * (size_t)(ms * ns > 0 ? ms * ns : 1));
    for (long z = 0; z < f.nbat; z++)
    {
        unsigned touched = 0;
        for (int p = 0; p < np; p++)
        {
            int first[4];
            for (int t = 0; t < P[p].C.n; t++)
            {
                first[t] = !(touched & (1u << P[p].C.sub[t]));
                touched |= 1u << P[p].C.sub[t];
            }
            leaf_product(&f, z, A, B, C, &P[p], d, ms, ns, ks, impl, alpha, first, beta, S, T, M);
        }
        if (!ref_fringe)
        {
            const long dm = d * ms, dn = d * ns, dk = d * ks;
            if (dm < f.m) classical_pass(&f, z, A, B, C, dm, f.m - dm, 0, f.n, 0, f.k, alpha, beta, 1);
            if (dn < f.n) classical_pass(&f, z, A, B, C, 0, dm, dn, f.n - dn, 0, f.k, alpha, beta, 1);
            if (dk < f.k) classical_pass(&f, z, A, B, C, 0, dm, 0, dn, dk, f.k - dk, alpha, beta, 0);
        }
    }
    free(S);
    free(T);
    free(M);
    flat_free(&f);
    return 0;
}

/* ---------------------------------------------------------------------------------------------- */
/* Blocked ABC: the reference's CPU loop nest, for timing the CPU baseline at full size            */
/* ---------------------------------------------------------------------------------------------- */
/*
 * oracle_tensor_mult_blocked restates the reference's ABC path at the level of its loop nest rather than
 * its definition: per product, the 5-loop partition NC / KC / MC / NR / MR (stra_partition,
 * src/stra_nodes/stra_partm.hpp:12-133; blocksizes of src/configs/sandybridge/config.hpp:30-35 with the
 * register tile widened to 8 x 6), packing that sums the product's 2^L sub-blocks of A and of B into the
 * packed panels (stra_pack_row_panel, src/stra_nodes/stra_packm.hpp:367-683), a register micro-kernel
 * over the packed panels (mul then add, as the AVX kernels of src/configs/sandybridge/bli_stra_*asm_d8x4.c)
 * and the multi-C update C_t += coeff_t * alpha * AB of every C target, beta at the first KC panel of the
 * first product that writes C_t (stra_gemm_micro_kernel, src/stra_nodes/stra_gemm_ukr.hpp:626-930).
 * The remainder of the 2^L split is computed by classical passes.  Only the timing of bench.py's CPU
 * baseline uses it; tests pin it to oracle_tensor_mult.
 */
#include <immintrin.h>

#define BL_MR 8
#define BL_NR 6
#define BL_MC 96
#define BL_KC 256
#define BL_NC 4096

/* acc (8 x 6, column-major) = sum over kc of a[k][0..7] * b[k][0..5]; a, b packed panels */
__attribute__((target("avx2"))) static void ukr_8x6(long kc, const double* a, const double* b, double* acc)
{
    __m256d c0[BL_NR], c1[BL_NR];
    for (int j = 0; j < BL_NR; j++) c0[j] = c1[j] = _mm256_setzero_pd();
    for (long p = 0; p < kc; p++)
    {
        const __m256d a0 = _mm256_loadu_pd(a + BL_MR * p), a1 = _mm256_loadu_pd(a + BL_MR * p + 4);
        for (int j = 0; j < BL_NR; j++)
        {
            const __m256d bj = _mm256_broadcast_sd(b + BL_NR * p + j);
            c0[j] = _mm256_add_pd(c0[j], _mm256_mul_pd(a0, bj));
            c1[j] = _mm256_add_pd(c1[j], _mm256_mul_pd(a1, bj));
        }
    }
    for (int j = 0; j < BL_NR; j++)
    {
        _mm256_storeu_pd(acc + BL_MR * j, c0[j]);
        _mm256_storeu_pd(acc + BL_MR * j + 4, c1[j]);
    }
}

/* One product of a Strassen level on batch entry z through the blocked loop nest. */
static void blocked_product(const flat_t* f, long z, const double* A, const double* B, double* C, const prod_t* P,
                            long d, long ms, long ns, long ks, double alpha, const int* first, double beta,
                            double* Bp)
{
    const double* Az = A + f->ba[z];
    const double* Bz = B + f->bb[z];
    double* Cz = C + f->bc[z];
    long ar[4], ac[4], br[4], bc[4], cr[4], cc[4];
    for (int t = 0; t < P->A.n; t++) ar[t] = (P->A.sub[t] / d) * ms, ac[t] = (P->A.sub[t] % d) * ks;
    for (int t = 0; t < P->B.n; t++) br[t] = (P->B.sub[t] / d) * ks, bc[t] = (P->B.sub[t] % d) * ns;
    for (int t = 0; t < P->C.n; t++) cr[t] = (P->C.sub[t] / d) * ms, cc[t] = (P->C.sub[t] % d) * ns;
    for (long jc = 0; jc < ns; jc += BL_NC)
    {
        const long nc = ns - jc < BL_NC ? ns - jc : BL_NC;
        const long nsl = (nc + BL_NR - 1) / BL_NR;
        for (long pc = 0; pc < ks; pc += BL_KC)
        {
            const long kc = ks - pc < BL_KC ? ks - pc : BL_KC;
#pragma omp parallel
            {
                /* the B panel: sum of the product's B sub-blocks, NR-wide slivers of [k][NR] */
#pragma omp for schedule(static)
                for (long sl = 0; sl < nsl; sl++)
                    for (long p = 0; p < kc; p++)
                        for (int jj = 0; jj < BL_NR; jj++)
                        {
                            const long j = jc + sl * BL_NR + jj;
                            double v = 0.0;
                            if (j < jc + nc)
                                for (int t = 0; t < P->B.n; t++)
                                    v += P->B.coef[t] * Bz[f->bk[br[t] + pc + p] + f->bn[bc[t] + j]];
                            Bp[(sl * kc + p) * BL_NR + jj] = v;
                        }
                double* Ap = (double*)malloc(sizeof(double) * BL_MC * BL_KC);
                double acc[BL_MR * BL_NR];
#pragma omp for schedule(dynamic, 1)
                for (long ic = 0; ic < ms; ic += BL_MC)
                {
                    const long mc = ms - ic < BL_MC ? ms - ic : BL_MC;
                    const long msl = (mc + BL_MR - 1) / BL_MR;
                    /* the A block: sum of the product's A sub-blocks, MR-tall slivers of [k][MR] */
                    for (long sl = 0; sl < msl; sl++)
                        for (long p = 0; p < kc; p++)
                            for (int ii = 0; ii < BL_MR; ii++)
                            {
                                const long i = ic + sl * BL_MR + ii;
                                double v = 0.0;
                                if (i < ic + mc)
                                    for (int t = 0; t < P->A.n; t++)
                                        v += P->A.coef[t] * Az[f->am[ar[t] + i] + f->ak[ac[t] + pc + p]];
                                Ap[(sl * kc + p) * BL_MR + ii] = v;
                            }
                    for (long jr = 0; jr < nsl; jr++)
                        for (long ir = 0; ir < msl; ir++)
                        {
                            ukr_8x6(kc, Ap + ir * kc * BL_MR, Bp + jr * kc * BL_NR, acc);
                            const long i0 = ic + ir * BL_MR, j0 = jc + jr * BL_NR;
                            const long mr = ic + mc - i0 < BL_MR ? ic + mc - i0 : BL_MR;
                            const long nr = jc + nc - j0 < BL_NR ? jc + nc - j0 : BL_NR;
                            for (int t = 0; t < P->C.n; t++)
                            {
                                const double cf = alpha * P->C.coef[t];
                                const int apply_beta = first[t] && pc == 0;
                                for (long j = 0; j < nr; j++)
                                {
                                    double* cj = Cz + f->cn[cc[t] + j0 + j];
                                    for (long i = 0; i < mr; i++)
                                    {
                                        double* cp = cj + f->cm[cr[t] + i0 + i];
                                        const double v = cf * acc[BL_MR * j + i];
                                        if (apply_beta) *cp = (beta == 0.0 ? 0.0 : beta * *cp) + v;
                                        else *cp += v;
                                    }
                                }
                            }
                        }
                }
                free(Ap);
            }
        }
    }
}

int oracle_tensor_mult_blocked(int level, int nthreads, int ndA, const long* lenA, const long* strideA,
                               const char* idxA, const double* A, int ndB, const long* lenB, const long* strideB,
                               const char* idxB, const double* B, int ndC, const long* lenC, const long* strideC,
                               const char* idxC, double* C, double alpha, double beta)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    flat_t f;
    groups_t G;
    memset(&f, 0, sizeof(f));
    if (flatten(ndA, lenA, strideA, idxA, ndB, lenB, strideB, idxB, ndC, lenC, strideC, idxC, &f, &G)) return -1;
    if (G.Ao.n || G.Bo.n || G.Co.n)
    {
        flat_free(&f);
        return -2; /* pure contractions only */
    }
    const long d = 1L << level, ms = f.m / d, ns = f.n / d, ks = f.k / d;
    const int core = ms > 0 && ns > 0 && ks > 0;
    prod_t P[49];
    const int np = core ? make_products(level, P) : 0;
    double* Bp = (double*)malloc(sizeof(double) * BL_KC * (BL_NC + BL_NR));
    for (long z = 0; z < f.nbat; z++)
    {
        unsigned touched = 0;
        for (int p = 0; p < np; p++)
        {
            int first[4];
            for (int t = 0; t < P[p].C.n; t++)
            {
                first[t] = !(touched & (1u << P[p].C.sub[t]));
                touched |= 1u << P[p].C.sub[t];
            }
            blocked_product(&f, z, A, B, C, &P[p], d, ms, ns, ks, alpha, first, beta, Bp);
        }
        if (!core)
            classical_pass(&f, z, A, B, C, 0, f.m, 0, f.n, 0, f.k, alpha, beta, 1);
        else
        {
            const long dm = d * ms, dn = d * ns, dk = d * ks;
            if (dm < f.m) classical_pass(&f, z, A, B, C, dm, f.m - dm, 0, f.n, 0, f.k, alpha, beta, 1);
            if (dn < f.n) classical_pass(&f, z, A, B, C, 0, dm, dn, f.n - dn, 0, f.k, alpha, beta, 1);
            if (dk < f.k) classical_pass(&f, z, A, B, C, 0, dm, 0, dn, dk, f.k - dk, alpha, beta, 0);
        }
    }
    free(Bp);
    flat_free(&f);
    return 0;
}

/* ---------------------------------------------------------------------------------------------- */
/* synthetic data (same generator as the device fill kernel)                                       */
/* ---------------------------------------------------------------------------------------------- */

static uint64_t splitmix64(uint64_t x)
{
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void oracle_fill_uniform(double* p, long n, unsigned long long seed, int emax)
{
    for (long i = 0; i < n; i++)
    {
        const uint64_t key = ((uint64_t)seed << 32) + (uint64_t)i;
        double v = (double)(splitmix64(key) >> 11) * 0x1.0p-53 * 2.0 - 1.0;
        if (emax > 0)
        {
            const uint64_t u = splitmix64(key ^ 0xA5A5A5A5A5A5A5A5ull);
            const int e = (int)(u % (uint64_t)(2 * emax + 1)) - emax;
            v = ldexp(v, e);
        }
        p[i] = v;
    }
}

int oracle_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
