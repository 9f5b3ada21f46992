"""General index classes on the GPU (the reference's classical fallback, stra_mult_blas,
src/stra_internal/3t/stra_mult.cxx:1311-1382, with the Strassen path for the inner contraction):
labels only in A or only in B are summed out, labels only in C receive the result by broadcast,
repeated labels address a diagonal, and contractions without K (outer product, weighting) run too.

Bars: small-integer inputs bit-exact against numpy.einsum at levels 0-2 (exact in FP64); uniform
inputs within 1e-13 (relative Frobenius) of the oracle's brute-force definition.
"""
import sys

import numpy as np
import pytest

import cases
import oracle
import tblis_strassen as ts

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

from test_gpu_parity import run_gpu  # noqa: E402


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


CASES = [
    dict(m=(6,), n=(5,), k=(7,), a_only=(3,)),
    dict(m=(6, 3), n=(5,), k=(7,), b_only=(4,), c_only=(2,)),
    dict(m=(40,), n=(36,), k=(50,), a_only=(2,), b_only=(3,), c_only=(2,)),
    dict(m=(24,), n=(20,), k=(30,), batch=(3,), a_only=(2,), c_only=(3,)),
    dict(m=(64,), n=(48,), c_only=(2,)),                      # outer product with broadcast
    dict(m=(33,), n=(17,)),                                   # outer product
    dict(m=(20,), n=(12,), batch=(5,)),                       # weighting (batch, no K)
    dict(m=(70, 2), n=(66,), k=(80,), a_only=(2, 2)),
]


def einsum_general(case, a, b, c, alpha, beta):
    A, B = cases.host_tensor(case["shape_A"], a), cases.host_tensor(case["shape_B"], b)
    C = cases.host_tensor(case["shape_C"], c)
    return cases.einsum_ref(case, A, B, C, alpha, beta)


@pytest.mark.parametrize("ci", range(len(CASES)))
@pytest.mark.parametrize("level,impl", [(0, "abc"), (1, "abc"), (2, "abc"), (1, "ab"), (1, "naive")])
def test_general_exact(ci, level, impl):
    case = cases.make_case(900 + ci, **CASES[ci])
    say("general exact", ci, level, impl, case["idx_A"], case["idx_B"], case["idx_C"])
    n = [int(np.prod(case[f"shape_{x}"])) for x in "ABC"]
    a, b, c = cases.small_ints(n[0], 1), cases.small_ints(n[1], 2), cases.small_ints(n[2], 3)
    got = run_gpu(case, a, b, c, 2.0, -1.0, level, impl)
    want = einsum_general(case, a, b, c, 2.0, -1.0)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("ci", range(len(CASES)))
def test_general_vs_oracle(ci):
    case = cases.make_case(950 + ci, **CASES[ci])
    say("general oracle", ci)
    n = [int(np.prod(case[f"shape_{x}"])) for x in "ABC"]
    a, b, c = cases.uniform(n[0], 4), cases.uniform(n[1], 5), cases.uniform(n[2], 6)
    got = run_gpu(case, a, b, c, 0.5, 0.25, 1, "abc")
    A, B = cases.host_tensor(case["shape_A"], a), cases.host_tensor(case["shape_B"], b)
    C = cases.host_tensor(case["shape_C"], c.copy())
    oracle.tensor_mult(0.5, A, case["idx_A"], B, case["idx_B"], 0.25, C, case["idx_C"], level=0, impl="abc")
    assert cases.rel_err(got, np.array(C)) <= 1e-13


def test_diagonal_labels():
    # C[i,j] = sum_k A[i,i,k] B[k,j,j]: repeated labels address diagonals
    say("diagonal labels")
    rng = np.random.default_rng(3)
    a = rng.integers(-3, 4, size=(9, 9, 11)).astype(np.float64)
    b = rng.integers(-3, 4, size=(11, 7, 7)).astype(np.float64)
    want = np.einsum("iik,kjj->ij", a, b)
    A = torch.from_numpy(np.asfortranarray(a)).cuda()
    B = torch.from_numpy(np.asfortranarray(b)).cuda()
    for level in (0, 1):
        C = torch.zeros(9, 7, dtype=torch.float64, device="cuda")
        ts.tensor_mult(1.0, A, "iik", B, "kjj", 0.0, C, "ij", level=level, flags=ts.FLAG_SYNC)
        assert np.array_equal(C.cpu().numpy(), want)


@pytest.mark.parametrize("impl", ["abc", "naive", "ab"])
@pytest.mark.parametrize("m,n,k,order", [(256, 256, 256, "F"), (101, 77, 55, "C"), (130, 64, 96, "F")])
def test_matrix_api_exact(impl, m, n, k, order):
    # stra_tblis_matrix_mult{,_naive,_ab}: C = alpha*A.B + beta*C, strided (column- or row-major) views
    say("matrix api", impl, m, n, k, order)
    rng = np.random.default_rng(m + n + k)
    a = rng.integers(-3, 4, size=(m, k)).astype(np.float64)
    b = rng.integers(-3, 4, size=(k, n)).astype(np.float64)
    c = rng.integers(-3, 4, size=(m, n)).astype(np.float64)
    conv = np.asfortranarray if order == "F" else np.ascontiguousarray
    A, B, C = (torch.from_numpy(conv(x)).cuda() for x in (a, b, c))
    if order == "F":
        A, B, C = (x.t().contiguous().t() for x in (A, B, C))
    ts.matrix_mult(2.0, A, B, -1.0, C, impl=impl)
    assert np.array_equal(C.cpu().numpy(), 2.0 * a @ b - c)


def test_matrix_api_errors():
    A = torch.zeros(4, 5, dtype=torch.float64, device="cuda")
    B = torch.zeros(4, 6, dtype=torch.float64, device="cuda")
    C = torch.zeros(4, 6, dtype=torch.float64, device="cuda")
    with pytest.raises(ts.MistraError):
        ts.matrix_mult(1.0, A, B, 0.0, C)


@pytest.mark.parametrize("alpha,empty", [(0.0, False), (1.0, True)])
def test_scale_only_touches_the_diagonal_of_c(alpha, empty):
    """alpha == 0 (or an empty operand) leaves C = beta*C over the elements C[idx_C] addresses: with a
    repeated label in idx_C that is a diagonal (the reference applies diagonal() to C before it scales,
    src/stra_iface/3t/stra_mult.cxx:47-50, :212-231); the elements off the diagonal keep their values."""
    L, K = 7, (0 if empty else 5)
    rng = np.random.default_rng(5)
    c = rng.uniform(-1, 1, size=L * L * 3)
    a = rng.uniform(-1, 1, size=max(1, L * K))[: L * K]
    b = rng.uniform(-1, 1, size=max(1, K * 3))[: K * 3]
    A, _ = dev_t((L, K), a)
    B, _ = dev_t((K, 3), b)
    C, cbuf = dev_t((L, L, 3), c)
    ts.tensor_mult(alpha, A, "ik", B, "kj", 0.5, C, "iij", level=1, flags=ts.FLAG_SYNC)
    got = cbuf.cpu().numpy().reshape((L, L, 3), order="F")
    want = c.reshape((L, L, 3), order="F").copy()
    for i in range(L):
        want[i, i, :] *= 0.5
    assert np.array_equal(got, want)


def dev_t(shape, flat):
    buf = torch.from_numpy(np.ascontiguousarray(flat, dtype=np.float64)).to("cuda")
    return buf.as_strided(tuple(shape), tuple(cases.colmajor_strides(shape))), buf
