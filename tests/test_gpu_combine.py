"""GPU: the one-pass combine of staged products into C against its per-sub-block form.

AB, Naive and the low-precision modes stage every leaf product M_p and add them into the C sub-blocks
afterwards.  The one-pass form reads each M_p once and updates every sub-block it feeds; the
per-sub-block form reads, for one sub-block at a time, the products of its term list.  Both add each
sub-block's terms in the same (table) order with the same +-1 coefficients, so their results must be
identical bit for bit, and both must agree with einsum within the FP64 bar.  The per-sub-block form is
forced through mistra_debug_set_combine_form.
"""
import ctypes
import sys

import numpy as np
import pytest

import cases
import tblis_strassen as ts

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def set_form(per_sub):
    lib = ts.load()
    lib.mistra_debug_set_combine_form.restype = ctypes.c_int
    lib.mistra_debug_set_combine_form.argtypes = [ctypes.c_int]
    return lib.mistra_debug_set_combine_form(per_sub)


def dev(shape, flat):
    buf = torch.from_numpy(np.ascontiguousarray(flat)).to(device="cuda", dtype=torch.float64)
    return buf.as_strided(tuple(shape), tuple(cases.colmajor_strides(shape))), buf


def run(case, a, b, c, alpha, beta, level, impl, precision):
    A, _ = dev(case["shape_A"], a)
    B, _ = dev(case["shape_B"], b)
    C, cbuf = dev(case["shape_C"], c)
    ts.tensor_mult(alpha, A, case["idx_A"], B, case["idx_B"], beta, C, case["idx_C"], level=level, impl=impl,
                   precision=precision, flags=ts.FLAG_SYNC)
    return cbuf.cpu().numpy()


SHAPES = [
    dict(m=(16, 16), n=(8, 16), k=(32, 4)),                # even split, permuted labels
    dict(m=(24, 7), n=(40,), k=(9, 6), batch=(3,)),        # fringes in m and k, a batch label
    dict(m=(130,), n=(66, 3), k=(70,)),                    # ragged: every fringe pass
]


@pytest.mark.parametrize("shape", range(len(SHAPES)))
@pytest.mark.parametrize("level,impl,precision", [
    (1, "ab", "native"), (2, "ab", "native"), (1, "naive", "native"), (2, "naive", "native"),
    (1, "naive", "bf16x3"), (2, "abc", "ozaki_int8"),
])
def test_one_pass_combine_equals_per_sub_block(shape, level, impl, precision):
    case = cases.make_case(700 + shape, **SHAPES[shape])
    n = [int(np.prod(case[f"shape_{x}"])) for x in "ABC"]
    a, b, c = cases.uniform(n[0], 31 + shape), cases.uniform(n[1], 32 + shape), cases.uniform(n[2], 33 + shape)
    alpha, beta = 0.75, -1.25
    say(f"combine forms: shape {SHAPES[shape]} L{level} {impl} {precision}")
    old = set_form(0)
    try:
        one = run(case, a, b, c, alpha, beta, level, impl, precision)
        set_form(1)
        per = run(case, a, b, c, alpha, beta, level, impl, precision)
    finally:
        set_form(old)
    assert np.array_equal(one, per)
    ref = cases.einsum_ref(case, cases.host_tensor(case["shape_A"], a), cases.host_tensor(case["shape_B"], b),
                           cases.host_tensor(case["shape_C"], c), alpha, beta)
    got = np.array(cases.host_tensor(case["shape_C"], one))
    tol = 1e-12 if precision == "native" else (1e-9 if precision == "ozaki_int8" else 1e-3)
    assert cases.rel_err(got, np.array(ref)) <= tol
