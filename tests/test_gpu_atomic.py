"""ABC with every sum staged and MISTRA_FLAG_ATOMIC_C: all products in one leaf launch, each adding
alpha*coeff*M_p into its C sub-blocks with FP64 atomic adds after one beta pass over C.

Checks, on the parity cases (fringes, multi-index groups, a batch label) and layouts that transpose the
plan:
  * small integers bit-exact against numpy.einsum (every partial sum is an exact integer, so the order
    of the atomic adds cannot show) at L0 / L1 / L2;
  * uniform FP64 within 1e-13 of the oracle, alpha/beta with beta = 0 over a NaN-filled C;
  * one leaf launch per call where the form applies (the kernel-timing hook counts them);
  * two runs of the same call within 1e-15 of each other (the adds into one element come in no fixed
    order, so the results may differ in the last bits, not more).
"""
import numpy as np
import pytest

import cases
import tblis_strassen as ts

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

from test_gpu_parity import CASES, dev, einsum, inputs, run_gpu, run_oracle, say  # noqa: E402

FORMS = {"L1 staged": (1, ts.FLAG_OUTER_STAGED), "L1 all": (1, ts.FLAG_ALL_STAGED),
         "L2 all": (2, ts.FLAG_ALL_STAGED)}


@pytest.mark.parametrize("ci", range(len(CASES)))
@pytest.mark.parametrize("form", sorted(FORMS))
def test_atomic_exact_on_integers(ci, form):
    level, flags = FORMS[form]
    case = cases.make_case(10 + ci, **CASES[ci])
    say(f"[atomic int] case={ci} {form}")
    a, b, c = inputs(case, 700 + ci, ints=True)
    got = run_gpu(case, a, b, c, 1.0, 1.0, level, "abc", flags=flags | ts.FLAG_ATOMIC_C)
    assert np.array_equal(got, einsum(case, a, b, c, 1.0, 1.0))


@pytest.mark.parametrize("ci", range(len(CASES)))
@pytest.mark.parametrize("form", sorted(FORMS))
def test_atomic_float_vs_oracle(ci, form):
    level, flags = FORMS[form]
    case = cases.make_case(20 + ci, **CASES[ci])
    say(f"[atomic f64] case={ci} {form}")
    a, b, c = inputs(case, 800 + ci)
    got = run_gpu(case, a, b, c, 1.0, 1.0, level, "abc", flags=flags | ts.FLAG_ATOMIC_C)
    ref = run_oracle(case, a, b, c, 1.0, 1.0, level=level, impl="abc")
    assert cases.rel_err(got, ref) <= 1e-13


@pytest.mark.parametrize("form", sorted(FORMS))
@pytest.mark.parametrize("alpha,beta", [(0.5, 0.0), (-2.0, 3.0), (1.0, -1.0), (0.75, 1.0)])
def test_atomic_alpha_beta(form, alpha, beta):
    level, flags = FORMS[form]
    case = cases.make_case(32, m=(8, 8), n=(16, 4), k=(4, 16))
    say(f"[atomic alpha/beta] {form} alpha={alpha} beta={beta}")
    a, b, c = inputs(case, 900, ints=True)
    if beta == 0.0:
        c = np.full_like(c, np.nan)
    got = run_gpu(case, a, b, c, alpha, beta, level, "abc", flags=flags | ts.FLAG_ATOMIC_C)
    assert np.array_equal(got, einsum(case, a, b, np.nan_to_num(c) if beta == 0.0 else c, alpha, beta))


@pytest.mark.parametrize("form", sorted(FORMS))
@pytest.mark.parametrize("layout", ["c_rowmajor", "a_kfast", "b_nfast", "strided"])
def test_atomic_layouts(layout, form):
    level, flags = FORMS[form]
    m, n, k = 200, 168, 136
    say(f"[atomic layout] {layout} {form}")
    a = cases.small_ints(m * k * 2, 21)
    b = cases.small_ints(k * n * 2, 22)
    c = cases.small_ints(m * n * 2, 23)
    sA, sB, sC = [1, m], [1, k], [1, m]
    if layout == "c_rowmajor":
        sC = [n, 1]
    elif layout == "a_kfast":
        sA = [k, 1]
    elif layout == "b_nfast":
        sB = [n, 1]
    else:
        sA, sB, sC = [2, 2 * m], [2, 2 * k], [2, 2 * m]
    A = torch.from_numpy(a).cuda().as_strided((m, k), sA)
    B = torch.from_numpy(b).cuda().as_strided((k, n), sB)
    cbuf = torch.from_numpy(c).cuda()
    C = cbuf.as_strided((m, n), sC)
    want = C.cpu().numpy() + A.cpu().numpy() @ B.cpu().numpy()
    ts.tensor_mult(1.0, A, "ik", B, "kj", 1.0, C, "ij", level=level, impl="abc",
                   flags=flags | ts.FLAG_ATOMIC_C | ts.FLAG_SYNC)
    assert np.array_equal(C.cpu().numpy(), want)


@pytest.mark.parametrize("form", sorted(FORMS))
def test_atomic_one_launch_and_run_to_run(form):
    level, flags = FORMS[form]
    case = cases.make_case(44, m=(96, 4), n=(48, 8), k=(128, 2))
    say(f"[atomic launches] {form}")
    n = [int(np.prod(case[f"shape_{x}"])) for x in "ABC"]
    A, _ = dev(case["shape_A"], cases.uniform(n[0], 1))
    B, _ = dev(case["shape_B"], cases.uniform(n[1], 2))
    outs = []
    for rep in range(2):
        C, cbuf = dev(case["shape_C"], np.zeros(n[2]))
        ts.kernel_timing(True)
        ts.kernel_timing_read(reset=True)
        ts.tensor_mult(1.0, A, case["idx_A"], B, case["idx_B"], 0.0, C, case["idx_C"], level=level, impl="abc",
                       flags=flags | ts.FLAG_ATOMIC_C | ts.FLAG_SYNC)
        _, launches, _ = ts.kernel_timing_read(reset=True)
        ts.kernel_timing(False)
        assert launches == 1, launches
        outs.append(cbuf.cpu().numpy())
    assert cases.rel_err(outs[0], outs[1]) <= 1e-15
