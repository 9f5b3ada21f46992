"""GPU parity of the HIP Strassen tensor contraction (through the C-ABI) against the oracle.

Bars (FP64 unless stated):
  * small integer inputs: bit-exact against numpy.einsum at every level and impl (Strassen over
    small integers is exact in FP64, so any wrong index, sign, sub-block or epilogue shows);
  * uniform[-1,1) inputs: relative Frobenius error <= 1e-13 against the oracle run with the same
    level and impl, and against einsum;
  * full-size configurations: L=1 within 1e-12 and L=2 within 5e-12 of the GPU L=0 result, which is
    itself pinned to the oracle at the 1024^3 configuration (within 1e-13);
  * FP32: within 1e-5 of the FP64 einsum (relative Frobenius).
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


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def dev(shape, flat, dtype=torch.float64):
    """Column-major device tensor of `shape` over a copy of the flat numpy array."""
    buf = torch.from_numpy(np.ascontiguousarray(flat)).to(device="cuda", dtype=dtype)
    return buf.as_strided(tuple(shape), tuple(cases.colmajor_strides(shape))), buf


def host_of(buf, shape):
    return cases.host_tensor(shape, buf.double().cpu().numpy())


def run_gpu(case, a, b, c, alpha, beta, level, impl, flags=0, dtype=torch.float64):
    A, _ = dev(case["shape_A"], a, dtype)
    B, _ = dev(case["shape_B"], b, dtype)
    C, cbuf = dev(case["shape_C"], c, dtype)
    ts.tensor_mult(alpha, A, case["idx_A"], B, case["idx_B"], beta, C, case["idx_C"], level=level, impl=impl,
                   flags=flags | ts.FLAG_SYNC)
    return np.array(host_of(cbuf, case["shape_C"]))


def run_oracle(case, a, b, c, alpha, beta, **kw):
    A, B = cases.host_tensor(case["shape_A"], a), cases.host_tensor(case["shape_B"], b)
    C = cases.host_tensor(case["shape_C"], c.copy())
    oracle.tensor_mult(alpha, A, case["idx_A"], B, case["idx_B"], beta, C, case["idx_C"], **kw)
    return np.array(C)


def inputs(case, seed, ints=False):
    n = [int(np.prod(case[f"shape_{x}"])) for x in "ABC"]
    gen = cases.small_ints if ints else cases.uniform
    return gen(n[0], seed), gen(n[1], seed + 1), gen(n[2], seed + 2)


def einsum(case, a, b, c, alpha, beta):
    return cases.einsum_ref(case, cases.host_tensor(case["shape_A"], a), cases.host_tensor(case["shape_B"], b),
                            cases.host_tensor(case["shape_C"], c), alpha, beta)


CASES = [
    dict(m=(64,), n=(64,), k=(64,)),
    dict(m=(8, 8), n=(8, 8), k=(8, 8)),
    dict(m=(4, 4, 8), n=(4, 4, 8), k=(4, 4, 8)),
    dict(m=(16, 16), n=(8, 16), k=(32, 4)),
    dict(m=(24,), n=(40,), k=(8,), batch=(3,)),
    dict(m=(160,), n=(136,), k=(200,)),
    dict(m=(12, 20), n=(36,), k=(6, 12)),
]
IMPLS = ["abc", "ab", "naive"]


def test_fill_and_error_kernels():
    x = torch.empty(10007, dtype=torch.float64, device="cuda")
    ts.fill_uniform(x, 42)
    torch.cuda.synchronize()
    assert np.array_equal(x.cpu().numpy(), cases.uniform(10007, 42))
    ts.fill_uniform(x, 3, emax=8)
    torch.cuda.synchronize()
    assert np.array_equal(x.cpu().numpy(), cases.uniform(10007, 3, 8))
    y = x + 1e-3 * torch.ones_like(x)
    want = cases.rel_err(y.cpu().numpy(), x.cpu().numpy())
    assert abs(ts.rel_error(y, x) - want) <= 1e-12 * want
    xf = torch.empty(999, dtype=torch.float32, device="cuda")
    ts.fill_uniform(xf, 9)
    torch.cuda.synchronize()
    assert np.array_equal(xf.cpu().numpy(), cases.uniform(999, 9).astype(np.float32))


@pytest.mark.parametrize("ci", range(len(CASES)))
@pytest.mark.parametrize("level", [0, 1, 2])
@pytest.mark.parametrize("impl", IMPLS)
def test_exact_on_integers(ci, level, impl):
    case = cases.make_case(10 + ci, **CASES[ci])
    say(f"[int] case={ci} {case['idx_C']}={case['idx_A']}*{case['idx_B']} L={level} {impl}")
    a, b, c = inputs(case, 100 + ci, ints=True)
    got = run_gpu(case, a, b, c, 1.0, 1.0, level, impl)
    assert np.array_equal(got, einsum(case, a, b, c, 1.0, 1.0))


@pytest.mark.parametrize("ci", range(len(CASES)))
@pytest.mark.parametrize("level", [0, 1, 2])
@pytest.mark.parametrize("impl", IMPLS)
def test_float_vs_oracle(ci, level, impl):
    case = cases.make_case(20 + ci, **CASES[ci])
    say(f"[f64] case={ci} L={level} {impl}")
    a, b, c = inputs(case, 200 + ci)
    got = run_gpu(case, a, b, c, 1.0, 1.0, level, impl)
    ref = run_oracle(case, a, b, c, 1.0, 1.0, level=level, impl=impl)
    assert cases.rel_err(got, ref) <= 1e-13
    assert cases.rel_err(got, einsum(case, a, b, c, 1.0, 1.0)) <= 1e-13


@pytest.mark.parametrize("level", [1, 2])
@pytest.mark.parametrize("impl", IMPLS)
@pytest.mark.parametrize("alpha,beta", [(0.5, 0.0), (-2.0, 3.0), (1.0, -1.0)])
def test_alpha_beta(level, impl, alpha, beta):
    case = cases.make_case(30 + level, m=(8, 8), n=(16, 4), k=(4, 16))
    say(f"[alpha/beta] L={level} {impl} alpha={alpha} beta={beta}")
    a, b, c = inputs(case, 300)
    if beta == 0.0:
        c = np.full_like(c, np.nan)  # beta == 0 must not read C
    got = run_gpu(case, a, b, c, alpha, beta, level, impl)
    want = einsum(case, a, b, np.nan_to_num(c) if beta == 0.0 else c, alpha, beta)
    assert cases.rel_err(got, want) <= 1e-13


def test_alpha_zero_scales_c():
    case = cases.make_case(40, m=(8,), n=(8,), k=(8,))
    a, b, c = inputs(case, 400)
    got = run_gpu(case, a, b, c, 0.0, 2.5, 1, "abc")
    assert np.array_equal(got, 2.5 * cases.host_tensor(case["shape_C"], c))
    got = run_gpu(case, a, b, c, 0.0, 0.0, 1, "abc")
    assert not got.any()


@pytest.mark.parametrize("level", [1, 2])
@pytest.mark.parametrize("impl", IMPLS)
def test_fringes(level, impl):
    case = cases.make_case(50 + level, m=(37,), n=(3, 9), k=(45,))
    say(f"[fringe] L={level} {impl}")
    a, b, c = inputs(case, 500, ints=True)
    assert np.array_equal(run_gpu(case, a, b, c, 1.0, 1.0, level, impl), einsum(case, a, b, c, 1.0, 1.0))
    # MISTRA_FLAG_REF_FRINGE reproduces the reference's truncation exactly
    got = run_gpu(case, a, b, c, 1.0, 1.0, level, impl, flags=ts.FLAG_REF_FRINGE)
    ref = run_oracle(case, a, b, c, 1.0, 1.0, level=level, impl=impl, ref_fringe=True)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("layout", ["c_rowmajor", "a_kfast", "b_nfast", "strided"])
@pytest.mark.parametrize("level", [0, 1, 2])
def test_layouts(layout, level):
    m, n, k = 96, 80, 72
    say(f"[layout] {layout} L={level}")
    a = cases.small_ints(m * k * 2, 1)
    b = cases.small_ints(k * n * 2, 2)
    c = cases.small_ints(m * n * 2, 3)
    sA, sB, sC = [1, m], [1, k], [1, m]
    if layout == "c_rowmajor":
        sC = [n, 1]
    elif layout == "a_kfast":
        sA = [k, 1]
    elif layout == "b_nfast":
        sB = [n, 1]  # column-major B is k-fast; this makes its n index unit-stride
    elif layout == "strided":
        sA, sB, sC = [2, 2 * m], [2, 2 * k], [2, 2 * m]
    ta = torch.from_numpy(a).cuda()
    tb = torch.from_numpy(b).cuda()
    tc = torch.from_numpy(c).cuda()
    A = ta.as_strided((m, k), sA)
    B = tb.as_strided((k, n), sB)
    C = tc.as_strided((m, n), sC)
    want = A.cpu().numpy() @ B.cpu().numpy() + C.cpu().numpy()
    ts.tensor_mult(1.0, A, "ik", B, "kj", 1.0, C, "ij", level=level, flags=ts.FLAG_SYNC)
    assert np.array_equal(C.cpu().numpy(), want)


def test_a_kfast_b_kfast_c_rowmajor():
    # A with k unit-stride, B with k unit-stride, C row-major (so the plan transposes)
    m, n, k = 130, 70, 90
    A = torch.from_numpy(cases.small_ints(m * k, 4)).cuda().as_strided((m, k), (k, 1))
    B = torch.from_numpy(cases.small_ints(k * n, 5)).cuda().as_strided((k, n), (1, k))
    C = torch.from_numpy(cases.small_ints(m * n, 6)).cuda().as_strided((m, n), (n, 1))
    want = A.cpu().numpy() @ B.cpu().numpy() + C.cpu().numpy()
    for level in (0, 1, 2):
        Cw = C.clone().as_strided((m, n), (n, 1))
        Cw.copy_(C)
        ts.tensor_mult(1.0, A, "ik", B, "kj", 1.0, Cw, "ij", level=level, flags=ts.FLAG_SYNC)
        assert np.array_equal(Cw.cpu().numpy(), want)


@pytest.mark.parametrize("level", [0, 1, 2])
@pytest.mark.parametrize("impl", IMPLS)
def test_float32(level, impl):
    case = cases.make_case(60, m=(16, 8), n=(32, 4), k=(8, 16))
    say(f"[f32] L={level} {impl}")
    a, b, c = inputs(case, 600)
    got = run_gpu(case, a, b, c, 1.0, 1.0, level, impl, dtype=torch.float32)
    want = einsum(case, a.astype(np.float32).astype(np.float64), b.astype(np.float32).astype(np.float64),
                  c.astype(np.float32).astype(np.float64), 1.0, 1.0)
    assert cases.rel_err(got, want) <= 1e-5


def test_dropin_symbols_device_and_host():
    case = cases.make_case(70, m=(8, 8), n=(8, 8), k=(8, 8))
    a, b, c = inputs(case, 700, ints=True)
    want = einsum(case, a, b, c, 1.0, 1.0)
    for fn, impl in ((ts.stra_mult, "abc"), (ts.stra_mult, "ab"), (ts.stra_mult_2level, "naive"), (ts.mult, "abc")):
        ts.set_impl(impl)
        # device tensors
        A, _ = dev(case["shape_A"], a)
        B, _ = dev(case["shape_B"], b)
        C, cbuf = dev(case["shape_C"], c)
        fn(1.0, A, case["idx_A"], B, case["idx_B"], 1.0, C, case["idx_C"])
        assert np.array_equal(np.array(host_of(cbuf, case["shape_C"])), want)
        # host arrays: staged through HBM
        Ah = cases.host_tensor(case["shape_A"], a.copy())
        Bh = cases.host_tensor(case["shape_B"], b.copy())
        ch = c.copy()
        Ch = cases.host_tensor(case["shape_C"], ch)
        fn(1.0, Ah, case["idx_A"], Bh, case["idx_B"], 1.0, Ch, case["idx_C"])
        assert np.array_equal(np.array(Ch), want)
    ts.set_impl("abc")


def test_errors_are_reported():
    A = torch.zeros(4, 4, dtype=torch.float64, device="cuda")
    B = torch.zeros(5, 4, dtype=torch.float64, device="cuda")
    C = torch.zeros(4, 4, dtype=torch.float64, device="cuda")
    with pytest.raises(ts.MistraError):
        ts.tensor_mult(1.0, A, "ik", B, "kj", 1.0, C, "ij", level=1)
    with pytest.raises(ts.MistraError):
        ts.stra_mult(1.0, A, "ik", B, "kj", 1.0, C, "ij")
    Bf = torch.zeros(4, 4, dtype=torch.float32, device="cuda")
    with pytest.raises(ts.MistraError):
        ts.tensor_mult(1.0, A, "ik", Bf, "kj", 1.0, C, "ij", level=1)


# ---- ABC with staged operand sums (MISTRA_FLAG_OUTER_STAGED, MISTRA_FLAG_ALL_STAGED) ---------------
# Same bars as the fused form: bit-exact on small integers, 1e-13 against the oracle's ABC on uniform
# inputs; layouts that transpose the plan, strided operands, fringes, alpha/beta with beta = 0 over a
# NaN-filled C.  At level 0 the flags pack A and B into K-fast copies.  The batched case (nbatch > 1)
# does not apply and runs fused.

@pytest.mark.parametrize("ci", range(len(CASES)))
@pytest.mark.parametrize("level", [0, 1, 2])
@pytest.mark.parametrize("flags", [ts.FLAG_OUTER_STAGED, ts.FLAG_ALL_STAGED, ts.FLAG_FUSED_SUMS])
def test_outer_staged_exact(ci, level, flags):
    case = cases.make_case(10 + ci, **CASES[ci])
    say(f"[outer int] case={ci} L={level} flags={flags}")
    a, b, c = inputs(case, 100 + ci, ints=True)
    got = run_gpu(case, a, b, c, 1.0, 1.0, level, "abc", flags=flags)
    assert np.array_equal(got, einsum(case, a, b, c, 1.0, 1.0))


@pytest.mark.parametrize("ci", range(len(CASES)))
@pytest.mark.parametrize("level", [1, 2])
@pytest.mark.parametrize("flags", [ts.FLAG_OUTER_STAGED, ts.FLAG_ALL_STAGED])
def test_outer_staged_float(ci, level, flags):
    case = cases.make_case(20 + ci, **CASES[ci])
    say(f"[outer f64] case={ci} L={level} flags={flags}")
    a, b, c = inputs(case, 200 + ci)
    got = run_gpu(case, a, b, c, 1.0, 1.0, level, "abc", flags=flags)
    ref = run_oracle(case, a, b, c, 1.0, 1.0, level=level, impl="abc")
    assert cases.rel_err(got, ref) <= 1e-13


@pytest.mark.parametrize("level", [1, 2])
@pytest.mark.parametrize("alpha,beta", [(0.5, 0.0), (-2.0, 3.0), (1.0, -1.0)])
@pytest.mark.parametrize("flags", [ts.FLAG_OUTER_STAGED, ts.FLAG_ALL_STAGED])
def test_outer_staged_alpha_beta(level, alpha, beta, flags):
    case = cases.make_case(32, m=(8, 8), n=(16, 4), k=(4, 16))
    say(f"[outer alpha/beta] L={level} alpha={alpha} beta={beta} flags={flags}")
    a, b, c = inputs(case, 300)
    if beta == 0.0:
        c = np.full_like(c, np.nan)
    got = run_gpu(case, a, b, c, alpha, beta, level, "abc", flags=flags)
    want = einsum(case, a, b, np.nan_to_num(c) if beta == 0.0 else c, alpha, beta)
    assert cases.rel_err(got, want) <= 1e-13


@pytest.mark.parametrize("level", [1, 2])
@pytest.mark.parametrize("flags", [ts.FLAG_OUTER_STAGED, ts.FLAG_ALL_STAGED])
def test_outer_staged_fringes(level, flags):
    case = cases.make_case(52, m=(37,), n=(3, 9), k=(45,))
    say(f"[outer fringe] L={level} flags={flags}")
    a, b, c = inputs(case, 500, ints=True)
    got = run_gpu(case, a, b, c, 1.0, 1.0, level, "abc", flags=flags)
    assert np.array_equal(got, einsum(case, a, b, c, 1.0, 1.0))


@pytest.mark.parametrize("level", [0, 1, 2])
@pytest.mark.parametrize("layout", ["c_rowmajor", "a_kfast", "b_nfast", "strided"])
@pytest.mark.parametrize("flags", [ts.FLAG_OUTER_STAGED, ts.FLAG_ALL_STAGED])
def test_outer_staged_layouts(layout, level, flags):
    m, n, k = 200, 168, 136
    say(f"[outer layout] {layout} L={level} flags={flags}")
    a = cases.small_ints(m * k * 2, 11)
    b = cases.small_ints(k * n * 2, 12)
    c = cases.small_ints(m * n * 2, 13)
    sA, sB, sC = [1, m], [1, k], [1, m]
    if layout == "c_rowmajor":
        sC = [n, 1]
    elif layout == "a_kfast":
        sA = [k, 1]
    elif layout == "b_nfast":
        sB = [n, 1]
    elif layout == "strided":
        sA, sB, sC = [2, 2 * m], [2, 2 * k], [2, 2 * m]
    A = torch.from_numpy(a).cuda().as_strided((m, k), sA)
    B = torch.from_numpy(b).cuda().as_strided((k, n), sB)
    C = torch.from_numpy(c).cuda().as_strided((m, n), sC)
    want = A.cpu().numpy() @ B.cpu().numpy() + C.cpu().numpy()
    ts.tensor_mult(1.0, A, "ik", B, "kj", 1.0, C, "ij", level=level, flags=ts.FLAG_SYNC | flags)
    assert np.array_equal(C.cpu().numpy(), want)
