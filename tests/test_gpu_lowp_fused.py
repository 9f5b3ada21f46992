"""Low-precision 1- and 2-level plans whose operands are K-fast form each operand's Strassen sums (7, or
49 through the Kronecker coefficients) and their digits in one launch.  The sums never reach HBM, but the
digits must be exactly those of the staged path.  That path runs the one-pass sums kernel, then the K-fast
slice kernel, and mistra_debug_set_lowp_fusion(0) forces it.

Checks, on shapes with padded rows and K, multi-index groups, a batch index, one operand K-fast and the
other not, and K offsets that are not affine (a K index after the M / N indices):
  * results equal to the staged path bit for bit (Ozaki BF16 with 2, 5 and 7 digits, Ozaki INT8 with 3
    and 6);
  * errors against FP64 numpy.einsum within the modes' bars: 1e-10 for the Ozaki modes at their default
    digits, 5e-4 for the 2-digit split.
"""
import numpy as np
import pytest

import cases
import tblis_strassen as ts

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

from test_gpu_precision import run, say  # noqa: E402


def kfast_case(seed, m, n, k, batch=(), b_kfast=True, k_split=False):
    """A case whose A (and, with b_kfast, B) has a K index first, i.e. unit-stride K (column-major).  With
    k_split the other K indices follow the M (N) indices, so the K offsets are not affine in the flat K
    position (the kernels then load them instead of computing them)."""
    base = cases.make_case(seed, m=m, n=n, k=k, batch=batch, permute=False)
    L = base["lengths"]
    labels = iter(sorted(L))
    ml = [next(labels) for _ in m]
    nl = [next(labels) for _ in n]
    kl = [next(labels) for _ in k]
    bl = [next(labels) for _ in batch]
    k1, k2 = (kl[:1], kl[1:]) if k_split else (kl, [])
    idx_A = "".join(k1 + ml + k2 + bl)
    idx_B = "".join((k1 + nl + k2) if b_kfast else (nl + kl)) + "".join(bl)
    idx_C = "".join(ml + nl + bl)
    return dict(idx_A=idx_A, idx_B=idx_B, idx_C=idx_C, lengths=L,
                **{f"shape_{x}": [L[c] for c in ix] for x, ix in zip("ABC", (idx_A, idx_B, idx_C))})


SHAPES = [
    dict(m=(96,), n=(80,), k=(200,)),
    dict(m=(16, 12), n=(10, 18), k=(8, 26)),
    dict(m=(256,), n=(130,), k=(64,), batch=(2,)),
    dict(m=(64,), n=(100,), k=(96,), b_kfast=False),
    dict(m=(40, 3), n=(72,), k=(12, 20), k_split=True),
]
MODES = [("ozaki_bf16", 5, 0), ("ozaki_bf16", 2, 0), ("ozaki_bf16", 7, 0), ("ozaki_int8", 6, 0),
         ("ozaki_int8", 3, 0)]
BAR = {("ozaki_bf16", 5): 1e-10, ("ozaki_bf16", 7): 1e-10, ("ozaki_int8", 6): 1e-10, ("ozaki_bf16", 2): 5e-4}


@pytest.fixture(autouse=True)
def _restore():
    yield
    ts.debug_set_lowp_fusion(True)
    ts.debug_set_lowp(0, 0, 0)


@pytest.mark.parametrize("level", [1, 2])
@pytest.mark.parametrize("si", range(len(SHAPES)))
@pytest.mark.parametrize("mi", range(len(MODES)))
def test_fused_sum_slices_equal_staged_path(si, mi, level):
    prec, nsl, _ = MODES[mi]
    shape = dict(SHAPES[si])
    b_kfast = shape.pop("b_kfast", True)
    k_split = shape.pop("k_split", False)
    case = kfast_case(1500 + si, b_kfast=b_kfast, k_split=k_split, **shape)
    say(f"fused sums+digits L{level}: {case['idx_C']} = {case['idx_A']} * {case['idx_B']}, {prec} with {nsl} digits")
    n = [int(np.prod(case[f"shape_{x}"])) for x in "ABC"]
    a, b, c = cases.uniform(n[0], 11, emax=8), cases.uniform(n[1], 12, emax=8), cases.uniform(n[2], 13)
    if prec == "ozaki_bf16":
        ts.debug_set_lowp(nsl, 0, 0)
    else:
        ts.debug_set_lowp(0, nsl, 0)
    ts.debug_set_lowp_fusion(True)
    fused = np.array(run(case, a, b, c, level, prec, 1.0, 0.5))
    ts.debug_set_lowp_fusion(False)
    staged = np.array(run(case, a, b, c, level, prec, 1.0, 0.5))
    assert np.array_equal(fused, staged)
    want = cases.einsum_ref(case, cases.host_tensor(case["shape_A"], a), cases.host_tensor(case["shape_B"], b),
                            cases.host_tensor(case["shape_C"], c), 1.0, 0.5)
    err = cases.rel_err(fused, want)
    say(f"  rel err vs FP64 einsum {err:.2e}")
    bar = BAR.get((prec, nsl))
    if bar is not None:
        assert err <= bar
