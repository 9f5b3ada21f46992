"""Pins the CPU oracle (oracle/oracle.c) before anything is checked against it.

Each check is chosen so that a plausible mistake fails it:
  * brute force vs numpy.einsum (a dropped index class, a wrong stride, alpha/beta misuse);
  * Strassen on small integers is exact, so L=1/L=2 must equal einsum bit for bit (a wrong sign,
    sub-block or coefficient changes the integers);
  * the Brent equations of the product tables (a table that is not a valid bilinear algorithm);
  * the tables against the reference's own lists, as data (order, ids and coefficients);
  * the flattened index layer against einsum through its scatter vectors (grouping/sort/fold);
  * the splitmix64 generator against a published test vector.
"""
import itertools
import json
import os

import numpy as np
import pytest

import cases
import oracle

HERE = os.path.dirname(os.path.abspath(__file__))

SMALL_CASES = [
    dict(m=(6,), n=(5,), k=(7,)),                   # a matrix product
    dict(m=(2, 3), n=(4,), k=(3, 2)),
    dict(m=(4, 4), n=(4, 4), k=(4, 4)),             # C1-like, scaled down
    dict(m=(2, 2, 4), n=(2, 2, 4), k=(2, 2, 4)),    # C2-like, scaled down
    dict(m=(3,), n=(4,), k=(2,), batch=(3,)),       # batch index
    dict(m=(3,), n=(), k=(5,)),                     # matrix-vector
    dict(m=(3,), n=(4,), k=()),                     # outer product
    dict(m=(3,), n=(2,), k=(4,), a_only=(2,), b_only=(3,), c_only=(2,)),  # general index classes
]


def _run_ref(case, A, B, C, alpha, beta):
    Ct = C.copy(order="F") if C.ndim else C.copy()
    Ch = cases.host_tensor(case["shape_C"], Ct.ravel(order="F")) if C.ndim else Ct
    oracle.contract_ref(alpha, A, case["idx_A"], B, case["idx_B"], beta, Ch, case["idx_C"])
    return Ch


def _inputs(case, seed, ints=False):
    nA = int(np.prod(case["shape_A"])) if case["shape_A"] else 1
    nB = int(np.prod(case["shape_B"])) if case["shape_B"] else 1
    nC = int(np.prod(case["shape_C"])) if case["shape_C"] else 1
    gen = cases.small_ints if ints else cases.uniform
    a, b, c = gen(nA, seed), gen(nB, seed + 1), gen(nC, seed + 2)
    return (cases.host_tensor(case["shape_A"], a), cases.host_tensor(case["shape_B"], b),
            cases.host_tensor(case["shape_C"], c))


@pytest.mark.parametrize("ci", range(len(SMALL_CASES)))
@pytest.mark.parametrize("alpha,beta", [(1.0, 1.0), (0.5, 0.0), (-2.0, 3.0)])
def test_contract_ref_vs_einsum(ci, alpha, beta):
    case = cases.make_case(100 + ci, **SMALL_CASES[ci])
    A, B, C = _inputs(case, 7 * ci)
    want = cases.einsum_ref(case, A, B, C, alpha, beta)
    got = _run_ref(case, A, B, C, alpha, beta)
    assert cases.rel_err(got, want) < 1e-14


@pytest.mark.parametrize("ci", range(len(SMALL_CASES)))
def test_contract_ref_exact_on_integers(ci):
    case = cases.make_case(200 + ci, **SMALL_CASES[ci])
    A, B, C = _inputs(case, 11 * ci, ints=True)
    want = cases.einsum_ref(case, A, B, C, 2.0, -1.0)
    got = _run_ref(case, A, B, C, 2.0, -1.0)
    assert np.array_equal(np.asarray(got), want)


def _run_oracle(case, A, B, C, alpha, beta, **kw):
    Cf = np.array(C, order="F", copy=True)
    Ch = cases.host_tensor(case["shape_C"], Cf.ravel(order="F"))
    oracle.tensor_mult(alpha, A, case["idx_A"], B, case["idx_B"], beta, Ch, case["idx_C"], **kw)
    return np.array(Ch)


STRASSEN_CASES = [
    dict(m=(4, 4), n=(4, 4), k=(4, 4)),
    dict(m=(2, 2, 4), n=(2, 2, 4), k=(2, 2, 4)),
    dict(m=(8,), n=(12,), k=(16,)),
    dict(m=(4,), n=(8,), k=(4,), batch=(3,)),
]


@pytest.mark.parametrize("ci", range(len(STRASSEN_CASES)))
@pytest.mark.parametrize("level", [0, 1, 2])
@pytest.mark.parametrize("impl", ["abc", "ab", "naive"])
def test_strassen_oracle_exact_on_integers(ci, level, impl):
    case = cases.make_case(300 + ci, **STRASSEN_CASES[ci])
    A, B, C = _inputs(case, 13 * ci, ints=True)
    want = cases.einsum_ref(case, A, B, C, 1.0, 1.0)
    got = _run_oracle(case, A, B, C, 1.0, 1.0, level=level, impl=impl)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("level", [1, 2])
@pytest.mark.parametrize("impl", ["abc", "ab", "naive"])
@pytest.mark.parametrize("alpha,beta", [(1.0, 1.0), (0.75, 0.0), (-1.5, 2.0)])
def test_strassen_oracle_float(level, impl, alpha, beta):
    case = cases.make_case(400 + level, m=(8, 4), n=(4, 8), k=(16, 2))
    A, B, C = _inputs(case, 17)
    want = cases.einsum_ref(case, A, B, C, alpha, beta)
    got = _run_oracle(case, A, B, C, alpha, beta, level=level, impl=impl)
    assert cases.rel_err(got, want) < 1e-14


@pytest.mark.parametrize("level", [1, 2])
def test_strassen_oracle_fringe(level):
    # m, n, k not divisible by 2^L: computed in full by default, truncated with ref_fringe
    case = cases.make_case(500 + level, m=(7,), n=(3, 3), k=(11,))
    A, B, C = _inputs(case, 23, ints=True)
    want = cases.einsum_ref(case, A, B, C, 1.0, 1.0)
    assert np.array_equal(_run_oracle(case, A, B, C, 1.0, 1.0, level=level), want)
    trunc = _run_oracle(case, A, B, C, 1.0, 1.0, level=level, ref_fringe=True)
    assert not np.array_equal(trunc, want)
    # the truncated result: only rows < d*floor(m/d), cols < d*floor(n/d) of the flattened C are
    # updated, with k summed over p < d*floor(k/d) (src/matrix/stra_tensor_view.hpp:246-247)
    d = 2 ** level
    (dims, vecs) = oracle.describe(case["shape_A"], cases.colmajor_strides(case["shape_A"]), case["idx_A"],
                                   case["shape_B"], cases.colmajor_strides(case["shape_B"]), case["idx_B"],
                                   case["shape_C"], cases.colmajor_strides(case["shape_C"]), case["idx_C"])
    m, n, k = dims[:3]
    am, ak, bk, bn, cm, cn = vecs
    dm, dn, dk = d * (m // d), d * (n // d), d * (k // d)
    a, b = np.asarray(A).ravel(order="F"), np.asarray(B).ravel(order="F")
    expect = np.asarray(C).ravel(order="F").copy()
    Af = a[am[:dm, None] + ak[None, :dk]]
    Bf = b[bk[:dk, None] + bn[None, :dn]]
    expect[(cm[:dm, None] + cn[None, :dn]).ravel()] += (Af @ Bf).ravel()
    assert np.array_equal(trunc.ravel(order="F"), expect)


# ---- tables ------------------------------------------------------------------------------------------

def _brent_ok(prods, d):
    """sum_r U[(i,j),r] V[(k,l),r] W[(p,q),r] == [j==k][i==p][l==q] over d x d block grids."""
    nb = d * d
    T = np.zeros((nb, nb, nb))
    for A, B, C in prods:
        u = np.zeros(nb)
        v = np.zeros(nb)
        w = np.zeros(nb)
        for s, c in A:
            u[s] += c
        for s, c in B:
            v[s] += c
        for s, c in C:
            w[s] += c
        T += np.einsum("a,b,c->abc", u, v, w)
    want = np.zeros_like(T)
    for i, j, l in itertools.product(range(d), repeat=3):
        want[i * d + j, j * d + l, i * d + l] = 1.0
    return np.array_equal(T, want)


@pytest.mark.parametrize("level", [1, 2])
def test_tables_satisfy_brent_equations(level):
    prods = oracle.products(level)
    assert len(prods) == 7 ** level
    assert _brent_ok(prods, 2 ** level)


def test_planted_table_mistake_is_caught():
    prods = [list(map(list, p)) for p in oracle.products(1)]
    prods[3][1][0] = (prods[3][1][0][0], -prods[3][1][0][1])  # flip one coefficient of B in M3
    assert not _brent_ok(prods, 2)
    prods = [list(map(list, p)) for p in oracle.products(2)]
    prods[20][2][1] = ((prods[20][2][1][0] + 1) % 16, prods[20][2][1][1])  # move one C target
    assert not _brent_ok(prods, 4)


def test_tables_match_reference_lists():
    with open(os.path.join(HERE, "golden", "strassen_tables.json")) as f:
        ref = json.load(f)
    for level, key in ((1, "level1"), (2, "level2")):
        mine = oracle.products(level)
        assert len(mine) == len(ref[key])
        for p, (A, B, C) in enumerate(mine):
            for op, terms in zip("ABC", (A, B, C)):
                assert [[s, c] for s, c in terms] == ref[key][p][op], (level, p, op)
    # U, V, W columns are the 1-level products
    U, V, W = (np.array(ref[x]) for x in "UVW")
    for p, (A, B, C) in enumerate(oracle.products(1)):
        for mat, terms in ((U, A), (V, B), (W, C)):
            col = np.zeros(4)
            for s, c in terms:
                col[s] = c
            assert np.array_equal(col, mat[:, p])


# ---- index layer -------------------------------------------------------------------------------------

@pytest.mark.parametrize("ci", range(7))
def test_flattening_matches_einsum(ci):
    case = cases.make_case(600 + ci, **SMALL_CASES[ci])
    sA, sB, sC = (cases.colmajor_strides(case[f"shape_{x}"]) for x in "ABC")
    (dims, vecs) = oracle.describe(case["shape_A"], sA, case["idx_A"], case["shape_B"], sB, case["idx_B"],
                                   case["shape_C"], sC, case["idx_C"])
    m, n, k, nb, pure = dims
    assert pure == 1
    am, ak, bk, bn, cm, cn = vecs
    assert (m, n, k) == cases.mnk(case)[:3] or nb > 1
    A, B, C = _inputs(case, 29)
    a, b = np.asarray(A).ravel(order="F"), np.asarray(B).ravel(order="F")
    want = cases.einsum_ref(case, A, B, np.zeros_like(np.asarray(C)), 1.0, 0.0).ravel(order="F")
    # one batch entry suffices when there is no batch index
    if nb == 1:
        Af = a[am[:, None] + ak[None, :]]
        Bf = b[bk[:, None] + bn[None, :]]
        got = np.zeros_like(want)
        got[(cm[:, None] + cn[None, :]).ravel()] = (Af @ Bf).ravel()
        assert cases.rel_err(got, want) < 1e-14
        # every element of A, B and C is reached exactly once
        assert sorted((am[:, None] + ak[None, :]).ravel()) == list(range(a.size))
        assert sorted((cm[:, None] + cn[None, :]).ravel()) == list(range(want.size))


def test_uniform_generator():
    # published first output of splitmix64 seeded with 0
    assert int(cases.splitmix64(np.array([0], dtype=np.uint64))[0]) == 0xE220A8397B1DCDAF
    for seed, emax in ((0, 0), (7, 0), (3, 8)):
        assert np.array_equal(oracle.fill_uniform(1000, seed, emax), cases.uniform(1000, seed, emax))
    v = cases.uniform(100000, 5)
    assert v.min() >= -1.0 and v.max() < 1.0 and abs(v.mean()) < 0.01


@pytest.mark.parametrize("groups", [dict(m=(16, 16), n=(8, 16), k=(32, 4)), dict(m=(13, 7), n=(11,), k=(9, 5), batch=(2,)),
                                    dict(m=(100,), n=(90,), k=(300,)), dict(m=(200,), n=(7, 9), k=(6, 50))])
@pytest.mark.parametrize("level", [0, 1, 2])
def test_blocked_port_equals_the_oracle(groups, level):
    """The blocked ABC port (the timed CPU baseline: packed panels of summed sub-blocks, 8 x 6 register
    tiles, fringes of the micro-tile, KC panels, multi-C updates) computes the oracle's contraction:
    bit for bit on small integers, within 1e-14 on uniform inputs, with alpha and beta."""
    case = cases.make_case(5 + level, **groups)
    n = [int(np.prod(case[f"shape_{x}"])) for x in "ABC"]
    for gen, tol in ((cases.small_ints, 0.0), (cases.uniform, 1e-14)):
        a, b, c = gen(n[0], 1), gen(n[1], 2), gen(n[2], 3)
        A, B = cases.host_tensor(case["shape_A"], a), cases.host_tensor(case["shape_B"], b)
        C1, C2 = cases.host_tensor(case["shape_C"], c.copy()), cases.host_tensor(case["shape_C"], c.copy())
        oracle.tensor_mult(2.0, A, case["idx_A"], B, case["idx_B"], -0.5, C1, case["idx_C"], level=level, impl="abc")
        oracle.tensor_mult_blocked(2.0, A, case["idx_A"], B, case["idx_B"], -0.5, C2, case["idx_C"], level=level)
        if tol == 0.0:
            assert np.array_equal(C1, C2)
        else:
            assert cases.rel_err(C2, C1) <= tol


def test_blocked_port_rejects_general_classes():
    case = cases.make_case(9, m=(6,), n=(5,), k=(7,), a_only=(3,))
    n = [int(np.prod(case[f"shape_{x}"])) for x in "ABC"]
    A, B = cases.host_tensor(case["shape_A"], np.ones(n[0])), cases.host_tensor(case["shape_B"], np.ones(n[1]))
    C = cases.host_tensor(case["shape_C"], np.zeros(n[2]))
    with pytest.raises(ValueError):
        oracle.tensor_mult_blocked(1.0, A, case["idx_A"], B, case["idx_B"], 0.0, C, case["idx_C"], level=1)
