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


FP8_EXACT = [  # leaf K-steps of 128 at L0: 1, 2, 3, 4, 5, 7 (the FP8 GEMM's loop and its tails)
    dict(m=(40, 3), n=(50,), k=(100,)),
    dict(m=(300,), n=(130,), k=(250,)),
    dict(m=(257,), n=(64,), k=(12, 25)),
    dict(m=(8, 32), n=(129,), k=(500,)),
    dict(m=(128,), n=(16, 9), k=(600,)),
    dict(m=(30,), n=(40,), k=(900,), batch=(2,)),
]


@pytest.mark.parametrize("level", [0, 1, 2])
@pytest.mark.parametrize("ci", range(len(FP8_EXACT)))
def test_fp8_exact_on_small_integers(ci, level):
    """Integers in [-3, 3]: every Strassen sum (|x| <= 12) has at most 4 significant bits, so its E4M3 code
    under the block's power-of-two scale is exact, the MFMA's products and FP32 sums stay exact (below
    2^24), and the FP32 products and their FP64 combine reproduce numpy.einsum bit for bit.  The shapes
    cover padded rows (M not a multiple of 256, odd M, whose products take the LDS-staged store), padded
    K and 1 to 7 K-steps per leaf."""
    case = cases.make_case(1400 + ci, **FP8_EXACT[ci])
    say("fp8 exact", ci, level)
    n = [int(np.prod(case[f"shape_{x}"])) for x in "ABC"]
    a, b, c = cases.small_ints(n[0], 40 + ci), cases.small_ints(n[1], 50 + ci), cases.small_ints(n[2], 60 + ci)
    got = run(case, a, b, c, level, "fp8_block", 2.0, -1.0)
    want = cases.einsum_ref(case, cases.host_tensor(case["shape_A"], a), cases.host_tensor(case["shape_B"], b),
                            cases.host_tensor(case["shape_C"], c), 2.0, -1.0)
    assert np.array_equal(got, want), cases.rel_err(got, want)


@pytest.mark.parametrize("level", [0, 1, 2])
@pytest.mark.parametrize("prec", ["ozaki_bf16", "ozaki_int8", "bf16x3", "fp8_block"])
def test_precision_on_float_tensors(level, prec):
    """FP32 tensors run through FP64 copies: the error against FP64 arithmetic on the same FP32 values is
    the mode's own (<= 1e-10 Ozaki, ...) plus one FP32 rounding of C (<= 1e-7 relative)."""
    case = cases.make_case(1250 + level, m=(12, 20), n=(36,), k=(6, 50))
    say("precision fp32", level, prec)
    n = [int(np.prod(case[f"shape_{x}"])) for x in "ABC"]
    a, b, c = (cases.uniform(n[i], 21 + i).astype(np.float32).astype(np.float64) for i in range(3))
    A, _ = dev(case["shape_A"], a, torch.float32)
    B, _ = dev(case["shape_B"], b, torch.float32)
    C, cbuf = dev(case["shape_C"], c, torch.float32)
    ts.tensor_mult(1.0, A, case["idx_A"], B, case["idx_B"], 0.5, C, case["idx_C"], level=level, impl="abc",
                   precision=prec, flags=ts.FLAG_SYNC)
    got = cases.host_tensor(case["shape_C"], cbuf.double().cpu().numpy())
    want = cases.einsum_ref(case, cases.host_tensor(case["shape_A"], a), cases.host_tensor(case["shape_B"], b),
                            cases.host_tensor(case["shape_C"], c), 1.0, 0.5)
    err = cases.rel_err(got, want)
    say("  rel err", err)
    bar = {"ozaki_bf16": 1e-7, "ozaki_int8": 1e-7, "bf16x3": 5e-4, "fp8_block": 0.05 * 2 ** level}[prec]
    assert err <= bar


def test_precision_float_strided_c_keeps_other_elements():
    """The FP64 copy of an FP32 C is written back only where C[idx_C] addresses it: the elements of the
    enclosing buffer between a strided view's elements keep their values."""
    base = torch.full((2 * 24 * 36,), 7.0, dtype=torch.float32, device="cuda")
    C = base.as_strided((24, 36), (2, 48))
    A = torch.rand(24, 40, dtype=torch.float32, device="cuda") - 0.5
    B = torch.rand(40, 36, dtype=torch.float32, device="cuda") - 0.5
    ts.tensor_mult(1.0, A, "ik", B, "kj", 0.0, C, "ij", level=1, precision="ozaki_int8", flags=ts.FLAG_SYNC)
    want = (A.double() @ B.double()).float()
    assert torch.allclose(C, want, rtol=1e-6, atol=1e-6)
    mask = torch.ones_like(base, dtype=torch.bool)
    mask.as_strided((24, 36), (2, 48)).fill_(False)
    assert torch.all(base[mask] == 7.0)


def test_precision_float_rejects_host_tensors():
    A = np.zeros((8, 8), dtype=np.float32)
    with pytest.raises(ts.MistraError):
        ts.tensor_mult(1.0, A, "ik", A, "kj", 0.0, A.copy(), "ij", level=1, precision="bf16x3")


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


@pytest.fixture
def lowp_defaults():
    yield
    ts.debug_set_lowp(0, 0, 0)


# K of 300 and 2 x 257 make several 64-deep K-steps, so a fold every 1 or 2 steps runs inside the loop
LOWP_VARIANT_CASES = [dict(m=(130,), n=(70,), k=(300,)), dict(m=(24, 5), n=(66,), k=(2, 257))]


@pytest.mark.parametrize("ci", range(len(LOWP_VARIANT_CASES)))
@pytest.mark.parametrize("level", [0, 1])
@pytest.mark.parametrize("prec,setting", [("ozaki_bf16", (6, 0, 0)), ("ozaki_bf16", (7, 0, 0)),
                                          ("ozaki_int8", (0, 6, 1)), ("ozaki_int8", (0, 6, 2))])
def test_ozaki_kernel_variants(ci, level, prec, setting, lowp_defaults):
    """The register-staged BF16 GEMM (6 or 7 digits) and the INT8 GEMM that folds its INT32 sums into
    FP64 inside the K loop (by default only for K beyond 65,536), forced on small shapes: within 1e-10 of
    FP64 on 2^u-scaled inputs and bit-exact on small integers."""
    ts.debug_set_lowp(*setting)
    case = cases.make_case(1500 + ci, **LOWP_VARIANT_CASES[ci])
    say("ozaki variant", ci, level, prec, setting)
    n = [int(np.prod(case[f"shape_{x}"])) for x in "ABC"]
    a, b, c = cases.uniform(n[0], 11, emax=8), cases.uniform(n[1], 12, emax=8), cases.uniform(n[2], 13)
    want = cases.einsum_ref(case, cases.host_tensor(case["shape_A"], a), cases.host_tensor(case["shape_B"], b),
                            cases.host_tensor(case["shape_C"], c), 1.0, 0.5)
    err = cases.rel_err(run(case, a, b, c, level, prec, 1.0, 0.5), want)
    say("  rel err", err)
    assert err <= 1e-10
    a, b, c = cases.small_ints(n[0], 14), cases.small_ints(n[1], 15), cases.small_ints(n[2], 16)
    got = run(case, a, b, c, level, prec, 2.0, -1.0)
    want = cases.einsum_ref(case, cases.host_tensor(case["shape_A"], a), cases.host_tensor(case["shape_B"], b),
                            cases.host_tensor(case["shape_C"], c), 2.0, -1.0)
    assert np.array_equal(got, want)


def test_debug_lowp_rejects_out_of_range():
    for bad in ((1, 0, 0), (8, 0, 0), (0, 7, 0), (0, 0, 1025), (0, 0, -1)):
        with pytest.raises(ts.MistraError):
            ts.debug_set_lowp(*bad)
