"""Low-precision leaf products with FP64 accuracy won back (MISTRA_PREC_*), checked against FP64.

Bars (relative Frobenius error vs the FP64 result; inputs uniform[-1,1) times 2^u, u uniform in
[-8, 8], as in the precision configuration):
  * Ozaki BF16 slicing (5 digits, 15 products): <= 1e-10 at L = 0, 1, 2;
  * Ozaki INT8 slicing (6 digits of 7 bits, 21 products): <= 1e-10 at L = 0, 1, 2;
  * BF16 hi/lo split (3 products): <= 5e-4;
  * FP8 E4M3 with per-128 K-block scales: <= 0.05 * 2^L (the Strassen operand sums of 2^L terms are
    quantized after summing, so the relative error doubles per level).
Small-integer inputs are exact in the digit slicing, so the BF16 and INT8 modes must reproduce
numpy.einsum bit for bit (every digit product and every FP32 chunk or INT32 sum is an exact integer).
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

from test_gpu_parity import dev  # noqa: E402

TOL = {"ozaki_bf16": 1e-10, "ozaki_int8": 1e-10, "bf16x3": 5e-4, "fp8_block": 0.05}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run(case, a, b, c, level, prec, alpha=1.0, beta=0.0):
    A, _ = dev(case["shape_A"], a)
    B, _ = dev(case["shape_B"], b)
    C, cbuf = dev(case["shape_C"], c)
    ts.tensor_mult(alpha, A, case["idx_A"], B, case["idx_B"], beta, C, case["idx_C"], level=level, impl="abc",
                   precision=prec, flags=ts.FLAG_SYNC)
    return cases.host_tensor(case["shape_C"], cbuf.cpu().numpy())


CASES = [
    dict(m=(64,), n=(64,), k=(64,)),
    dict(m=(12, 20), n=(36,), k=(6, 50)),
    dict(m=(130,), n=(70,), k=(300,), batch=(2,)),
    dict(m=(32, 32), n=(32, 32), k=(16, 16)),
]


@pytest.mark.parametrize("ci", range(len(CASES)))
@pytest.mark.parametrize("level", [0, 1, 2])
@pytest.mark.parametrize("prec", ["ozaki_bf16", "ozaki_int8", "bf16x3", "fp8_block"])
def test_precision_error(ci, level, prec):
    case = cases.make_case(1200 + ci, **CASES[ci])
    say("precision", ci, level, prec)
    n = [int(np.prod(case[f"shape_{x}"])) for x in "ABC"]
    a, b, c = cases.uniform(n[0], 1, emax=8), cases.uniform(n[1], 2, emax=8), cases.uniform(n[2], 3)
    got = run(case, a, b, c, level, prec, 1.0, 0.5)
    want = cases.einsum_ref(case, cases.host_tensor(case["shape_A"], a), cases.host_tensor(case["shape_B"], b),
                            cases.host_tensor(case["shape_C"], c), 1.0, 0.5)
    err = cases.rel_err(got, want)
    say("  rel err", err)
    assert err <= TOL[prec] * (2 ** level if prec == "fp8_block" else 1)


@pytest.mark.parametrize("level", [0, 1, 2])
@pytest.mark.parametrize("prec", ["ozaki_bf16", "ozaki_int8", "bf16x3"])
def test_bf16_modes_exact_on_small_integers(level, prec):
    case = cases.make_case(1300 + level, m=(40, 3), n=(50,), k=(70,))
    n = [int(np.prod(case[f"shape_{x}"])) for x in "ABC"]
    a, b, c = cases.small_ints(n[0], 4), cases.small_ints(n[1], 5), cases.small_ints(n[2], 6)
    got = run(case, a, b, c, level, prec, 2.0, -1.0)
    want = cases.einsum_ref(case, cases.host_tensor(case["shape_A"], a), cases.host_tensor(case["shape_B"], b),
                            cases.host_tensor(case["shape_C"], c), 2.0, -1.0)
    assert np.array_equal(got, want)


def test_precision_rejects_float():
    A = torch.zeros(8, 8, dtype=torch.float32, device="cuda")
    with pytest.raises(ts.MistraError):
        ts.tensor_mult(1.0, A, "ik", A, "kj", 0.0, A.clone(), "ij", level=1, precision="bf16x3")


def test_int8_digits_reach_the_last_bits():
    # the 6 INT8 digits (about 41 bits) must beat the 3-product BF16 split by orders of magnitude and
    # carry the value down to well below 1e-10 on values of very different magnitude in one row
    case = cases.make_case(1400, m=(96,), n=(80,), k=(257,))
    n = [int(np.prod(case[f"shape_{x}"])) for x in "ABC"]
    a, b, c = cases.uniform(n[0], 7, emax=20), cases.uniform(n[1], 8, emax=20), cases.uniform(n[2], 9)
    want = cases.einsum_ref(case, cases.host_tensor(case["shape_A"], a), cases.host_tensor(case["shape_B"], b),
                            cases.host_tensor(case["shape_C"], c), 1.0, 0.0)
    e8 = cases.rel_err(run(case, a, b, c, 0, "ozaki_int8"), want)
    e3 = cases.rel_err(run(case, a, b, c, 0, "bf16x3"), want)
    say("int8 vs bf16x3", e8, e3)
    assert e8 <= 1e-10 and e8 * 1e4 < e3
