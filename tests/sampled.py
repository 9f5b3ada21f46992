"""Sampled FP64 check of a full-size contraction: for a seeded set of C elements, the exact definition
C[c_m[i] + c_n[j]] = sum_p A[a_m[i] + a_k[p]] * B[b_k[p] + b_n[j]] is evaluated on the host in FP64
over the scatter vectors of the oracle's index layer (oracle_describe: grouping, stride ordering, fold
and flattening restated from src/stra_iface/3t/stra_mult.cxx:35-180 and src/util/tensor.hpp), so a
full-size result is pinned to the oracle without running the oracle's whole contraction.
"""
import numpy as np

import cases
import oracle


def describe_case(case):
    """(m, n, k) and the six scatter vectors of a dense column-major case (no batch indices)."""
    sh = [case[f"shape_{x}"] for x in "ABC"]
    st = [cases.colmajor_strides(s) for s in sh]
    m, n, k = cases.mnk(case)
    d, vecs = oracle.describe(sh[0], st[0], case["idx_A"], sh[1], st[1], case["idx_B"], sh[2], st[2],
                              case["idx_C"], cap=max(m, n, k))
    assert d[4] == 1 and d[3] == 1, "pure contraction without batch indices"
    return (d[0], d[1], d[2]), vecs


def sample(case, a_flat, b_flat, nsamp=4096, seed=0, chunk=256):
    """C offsets (column-major element offsets) of nsamp seeded C elements and their FP64 values of
    A.B (alpha = 1, beta = 0)."""
    (m, n, k), (am, ak, bk, bn, cm, cn) = describe_case(case)
    rng = np.random.default_rng(seed)
    I = rng.integers(0, m, nsamp)
    J = rng.integers(0, n, nsamp)
    a = np.asarray(a_flat, dtype=np.float64)
    b = np.asarray(b_flat, dtype=np.float64)
    ref = np.empty(nsamp)
    for s0 in range(0, nsamp, chunk):
        i, j = I[s0:s0 + chunk], J[s0:s0 + chunk]
        ra = a[am[i][:, None] + ak[None, :]]
        rb = b[bk[None, :] + bn[j][:, None]]
        ref[s0:s0 + chunk] = np.einsum("sk,sk->s", ra, rb)
    return cm[I] + cn[J], ref


def rel_err(got, ref):
    got = np.asarray(got, dtype=np.float64)
    return float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
