"""Adjoints dA = dC.B and dB = A.dC (mistra_tensor_mult_backward) on the GPU.

Pins (FP64):
  * the adjoint identity <dC, C(V, B)> = <dA, V> and <dC, C(A, W)> = <dB, W> for random directions
    V, W: the contraction is linear in each operand, so a central finite difference is exact and the
    identity holds to rounding (relative 1e-12), at every level and impl;
  * small-integer inputs: dA and dB bit-exact against numpy.einsum (Strassen over small integers is
    exact in FP64);
  * torch.autograd.gradcheck of the test-only autograd Function at index lengths 3 and 6.
FP32 (the 64^6 configuration, L=1 ABC): forward, dA and dB within 1e-4 of the FP64 classical results.
"""
import sys

import numpy as np
import pytest

import cases
import tblis_strassen as ts

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

from strassen_autograd import contract  # noqa: E402


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def dev(shape, flat, dtype=torch.float64):
    buf = torch.from_numpy(np.ascontiguousarray(flat)).to(device="cuda", dtype=dtype)
    return buf.as_strided(tuple(shape), tuple(cases.colmajor_strides(shape))), buf


def fwd(case, A, B, level, impl, dtype=torch.float64):
    n = int(np.prod(case["shape_C"]))
    C, cbuf = dev(case["shape_C"], np.zeros(n), dtype)
    ts.tensor_mult(1.0, A, case["idx_A"], B, case["idx_B"], 0.0, C, case["idx_C"], level=level, impl=impl,
                   flags=ts.FLAG_SYNC)
    return cbuf


def bwd(case, A, B, dC, level, impl, dtype=torch.float64):
    nA, nB = (int(np.prod(case[f"shape_{x}"])) for x in "AB")
    dA, abuf = dev(case["shape_A"], np.zeros(nA), dtype)
    dB, bbuf = dev(case["shape_B"], np.zeros(nB), dtype)
    ts.tensor_mult_backward(1.0, A, case["idx_A"], B, case["idx_B"], dC, case["idx_C"], dA=dA, dB=dB,
                            level=level, impl=impl)
    torch.cuda.synchronize()
    return abuf, bbuf


CASES = [
    dict(m=(8, 6), n=(10,), k=(4, 8)),
    dict(m=(16, 4), n=(8, 8), k=(12, 4)),
    dict(m=(9,), n=(7, 3), k=(10,), batch=(3,)),
    dict(m=(32, 32), n=(32, 32), k=(32, 32)),
]


@pytest.mark.parametrize("ci", range(len(CASES)))
@pytest.mark.parametrize("level,impl", [(0, "abc"), (1, "abc"), (2, "abc"), (1, "ab"), (2, "naive")])
def test_adjoint_identity(ci, level, impl):
    case = cases.make_case(500 + ci, **CASES[ci])
    say("adjoint identity", ci, level, impl, case["idx_A"], case["idx_B"], case["idx_C"])
    nA, nB, nC = (int(np.prod(case[f"shape_{x}"])) for x in "ABC")
    A, _ = dev(case["shape_A"], cases.uniform(nA, 1))
    B, _ = dev(case["shape_B"], cases.uniform(nB, 2))
    dC, dcbuf = dev(case["shape_C"], cases.uniform(nC, 3))
    V, vbuf = dev(case["shape_A"], cases.uniform(nA, 4))
    W, wbuf = dev(case["shape_B"], cases.uniform(nB, 5))
    dA, dB = bwd(case, A, B, dC, level, impl)
    cv = fwd(case, V, B, 0, "abc")  # C(V, B): linear in the first operand
    cw = fwd(case, A, W, 0, "abc")
    lhs_a, rhs_a = float(torch.dot(dcbuf, cv)), float(torch.dot(dA, vbuf))
    lhs_b, rhs_b = float(torch.dot(dcbuf, cw)), float(torch.dot(dB, wbuf))
    scale_a = float(torch.linalg.norm(dA) * torch.linalg.norm(vbuf))
    scale_b = float(torch.linalg.norm(dB) * torch.linalg.norm(wbuf))
    assert abs(lhs_a - rhs_a) <= 1e-12 * scale_a, (lhs_a, rhs_a)
    assert abs(lhs_b - rhs_b) <= 1e-12 * scale_b, (lhs_b, rhs_b)


@pytest.mark.parametrize("ci", range(3))
@pytest.mark.parametrize("level,impl", [(1, "abc"), (2, "abc"), (1, "naive")])
def test_adjoint_exact_small_ints(ci, level, impl):
    case = cases.make_case(600 + ci, **CASES[ci])
    say("adjoint ints", ci, level, impl)
    nA, nB, nC = (int(np.prod(case[f"shape_{x}"])) for x in "ABC")
    a, b, dc = cases.small_ints(nA, 11), cases.small_ints(nB, 12), cases.small_ints(nC, 13)
    A, _ = dev(case["shape_A"], a)
    B, _ = dev(case["shape_B"], b)
    dC, _ = dev(case["shape_C"], dc)
    dA, dB = bwd(case, A, B, dC, level, impl)
    Ah, Bh, dCh = (cases.host_tensor(case[f"shape_{x}"], v) for x, v in zip("ABC", (a, b, dc)))
    lab = {x: case[f"idx_{x}"] for x in "ABC"}
    want_a = np.einsum(f"{lab['C']},{lab['B']}->{lab['A']}", dCh, Bh)
    want_b = np.einsum(f"{lab['A']},{lab['C']}->{lab['B']}", Ah, dCh)
    got_a = cases.host_tensor(case["shape_A"], dA.cpu().numpy())
    got_b = cases.host_tensor(case["shape_B"], dB.cpu().numpy())
    assert np.array_equal(got_a, want_a)
    assert np.array_equal(got_b, want_b)


@pytest.mark.parametrize("length,fast", [(3, False), (6, True)])
def test_gradcheck(length, fast):
    # C[a,b,c,e] = sum_{d,f} A[d,c,a,f] B[d,f,b,e]
    say("gradcheck", length, fast)
    g = torch.Generator(device="cpu").manual_seed(7)
    A = torch.rand((length,) * 4, generator=g, dtype=torch.float64).cuda().requires_grad_()
    B = torch.rand((length,) * 4, generator=g, dtype=torch.float64).cuda().requires_grad_()
    for level in (0, 1):
        f = lambda a, b: contract(a, b, "dcaf", "dfbe", "abce", (length,) * 4, level=level)  # noqa: E731
        assert torch.autograd.gradcheck(f, (A, B), eps=1e-6, atol=1e-9, rtol=1e-7, fast_mode=fast)


def test_c5_fp32_forward_and_adjoints():
    # 64^6: C[a,b,c,e] = sum_{d,f} A[d,c,a,f] B[d,f,b,e], m = n = k = 4096, FP32 L=1 ABC vs FP64 L=0
    say("C5 FP32 64^6")
    L = 64
    n = L ** 4
    case = dict(idx_A="dcaf", idx_B="dfbe", idx_C="abce", shape_A=[L] * 4, shape_B=[L] * 4, shape_C=[L] * 4)
    errs = {}
    res = {}
    for dt, level in ((torch.float64, 0), (torch.float32, 1)):
        A = torch.empty(n, dtype=dt, device="cuda")
        B = torch.empty(n, dtype=dt, device="cuda")
        dC = torch.empty(n, dtype=dt, device="cuda")
        ts.fill_uniform(A, 1)
        ts.fill_uniform(B, 2)
        ts.fill_uniform(dC, 3)
        v = lambda x: x.view([L] * 4).as_strided([L] * 4, cases.colmajor_strides([L] * 4))  # noqa: E731
        C = fwd(case, v(A), v(B), level, "abc", dt)
        dA, dB = bwd(case, v(A), v(B), v(dC), level, "abc", dt)
        res[level] = (C.double(), dA.double(), dB.double())
    for name, x, y in zip(("C", "dA", "dB"), res[1], res[0]):
        errs[name] = float(torch.linalg.norm(x - y) / torch.linalg.norm(y))
    say("C5 errors", errs)
    assert all(e <= 1e-4 for e in errs.values()), errs
