"""Shared helpers of the tests: seeded contraction cases, the splitmix64 generator, layouts.

A case follows the reference's random_contraction (test/test_strassen.cxx:129-286): labels are
assigned m-group first, then n, then k (then batch / A-only / B-only / C-only), each operand's index
order is a seeded permutation, and tensors are dense column-major in that order.
"""
import numpy as np

MASK64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x):
    """Vectorised splitmix64 over uint64 (same constants as the device fill kernel)."""
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform(n, seed, emax=0):
    """uniform[-1,1) values of the library's generator: key = seed*2^32 + i."""
    with np.errstate(over="ignore"):
        key = (np.uint64(seed) << np.uint64(32)) + np.arange(n, dtype=np.uint64)
    v = (splitmix64(key) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53 * 2.0 - 1.0
    if emax > 0:
        u = splitmix64(key ^ np.uint64(0xA5A5A5A5A5A5A5A5))
        e = (u % np.uint64(2 * emax + 1)).astype(np.int64) - emax
        v = np.ldexp(v, e)
    return v


def small_ints(n, seed, lo=-3, hi=3):
    """Small integers as float64: every Strassen and classical sum over them is exact in FP64."""
    rng = np.random.default_rng(seed)
    return rng.integers(lo, hi + 1, size=n).astype(np.float64)


def colmajor_strides(shape):
    s, out = 1, []
    for L in shape:
        out.append(s)
        s *= L
    return out


def make_case(seed, m=(), n=(), k=(), batch=(), a_only=(), b_only=(), c_only=(), permute=True):
    """Return dict(idx_A, idx_B, idx_C, shape_A, shape_B, shape_C, lengths)."""
    rng = np.random.default_rng(seed)
    label = iter("abcdefghijklmnopqrstuvwxyz")
    idx_A, idx_B, idx_C, lengths = [], [], [], {}
    for group, ops in ((m, "AC"), (n, "BC"), (k, "AB"), (batch, "ABC"), (a_only, "A"), (b_only, "B"),
                       (c_only, "C")):
        for L in group:
            c = next(label)
            lengths[c] = int(L)
            if "A" in ops:
                idx_A.append(c)
            if "B" in ops:
                idx_B.append(c)
            if "C" in ops:
                idx_C.append(c)
    if permute:
        idx_A = [idx_A[i] for i in rng.permutation(len(idx_A))]
        idx_B = [idx_B[i] for i in rng.permutation(len(idx_B))]
        idx_C = [idx_C[i] for i in rng.permutation(len(idx_C))]
    return dict(idx_A="".join(idx_A), idx_B="".join(idx_B), idx_C="".join(idx_C),
                shape_A=[lengths[c] for c in idx_A], shape_B=[lengths[c] for c in idx_B],
                shape_C=[lengths[c] for c in idx_C], lengths=lengths)


def host_tensor(shape, values):
    """Column-major numpy view of `values` (a flat float64 array) with the given shape."""
    itemsize = values.itemsize
    return np.lib.stride_tricks.as_strided(values, shape=tuple(shape),
                                           strides=tuple(s * itemsize for s in colmajor_strides(shape)))


def flops(case):
    L = case["lengths"]
    tot = 1
    for c in L:
        tot *= L[c]
    return 2 * tot


def mnk(case):
    L = case["lengths"]
    A, B, C = set(case["idx_A"]), set(case["idx_B"]), set(case["idx_C"])
    m = int(np.prod([L[c] for c in A & C - B])) if A & C - B else 1
    n = int(np.prod([L[c] for c in B & C - A])) if B & C - A else 1
    k = int(np.prod([L[c] for c in A & B - C])) if A & B - C else 1
    return m, n, k


def einsum_ref(case, A, B, C, alpha=1.0, beta=1.0):
    """alpha * einsum(A,B) + beta * C in float64 (numpy's einsum, no optimisation)."""
    ab = set(case["idx_A"]) | set(case["idx_B"])
    out = "".join(c for c in case["idx_C"] if c in ab)  # C-only labels are broadcast afterwards
    r = np.einsum(f"{case['idx_A']},{case['idx_B']}->{out}", A, B, optimize=False)
    shape = [case["lengths"][c] if c in ab else 1 for c in case["idx_C"]]
    r = np.broadcast_to(r.reshape(shape), [case["lengths"][c] for c in case["idx_C"]])
    return alpha * r + beta * np.asarray(C)


def rel_err(x, y):
    x = np.asarray(x, dtype=np.float64).ravel()
    y = np.asarray(y, dtype=np.float64).ravel()
    den = np.linalg.norm(y)
    return float(np.linalg.norm(x - y) / (den if den > 0 else 1.0))


def c4_batch(seed, count=256, lo=64, hi=4096, lmin=8, lmax=72):
    """Random batch of contractions (the batched configuration): per entry 3..6 indices split among
    m/n/k (each >= 1), index lengths in [lmin, lmax], group sizes log-uniform in [lo, hi], seeded
    per-tensor index permutations.

    The size and length constraints conflict (a group of c indices spans only [lmin^c, lmax^c]), so a
    group's log-uniform target is clamped to that range and factored into c lengths: the first c-1 are
    the rounded c-th root of the target, the last absorbs the remainder; every length is clamped to
    [lmin, lmax]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        total = int(rng.integers(3, 7))
        counts = [1, 1, 1]
        for _ in range(total - 3):
            counts[int(rng.integers(0, 3))] += 1
        groups = []
        for c in counts:
            target = float(np.exp(rng.uniform(np.log(lo), np.log(hi))))
            target = min(max(target, float(lmin) ** c), float(lmax) ** c)
            root = int(round(target ** (1.0 / c)))
            lens = [min(max(root, lmin), lmax) for _ in range(c - 1)]
            rest = target / float(np.prod(lens)) if lens else target
            lens.append(min(max(int(round(rest)), lmin), lmax))
            groups.append(tuple(lens))
        out.append(make_case(int(rng.integers(0, 2 ** 31)), m=groups[0], n=groups[1], k=groups[2]))
    return out
