"""GPU: the allocation-free entry point (mistra_tensor_mult_noalloc) against mistra_tensor_mult.

Both run the same kernels in the same order on the same plan tables, so every result must be equal bit
for bit; the plan tables and staged workspaces of the allocation-free calls come from the library's fixed
scratch block, which keeps the previous call's contents (so two calls in a row, the second smaller, must
still agree).  Errors: staged products that do not fit the block, host tensors, low precision.
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

FORMS = {"": 0, "staged": ts.FLAG_OUTER_STAGED, "staged_all": ts.FLAG_ALL_STAGED, "fused": ts.FLAG_FUSED_SUMS}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def dev(shape, flat, dtype=torch.float64):
    buf = torch.from_numpy(np.ascontiguousarray(flat)).to(device="cuda", dtype=dtype)
    return buf.as_strided(tuple(shape), tuple(cases.colmajor_strides(shape))), buf


def both(case, level, impl, form, seed, dtype=torch.float64, alpha=1.5, beta=-0.5):
    """(no-alloc result, regular result) of one contraction on the same seeded inputs."""
    n = [int(np.prod(case[f"shape_{x}"])) for x in "ABC"]
    a, b, c = cases.uniform(n[0], seed), cases.uniform(n[1], seed + 1), cases.uniform(n[2], seed + 2)
    A, _ = dev(case["shape_A"], a, dtype)
    B, _ = dev(case["shape_B"], b, dtype)
    out = []
    for fn in (ts.tensor_mult_noalloc, ts.tensor_mult):
        C, cbuf = dev(case["shape_C"], c, dtype)
        fn(alpha, A, case["idx_A"], B, case["idx_B"], beta, C, case["idx_C"], level=level, impl=impl,
           flags=FORMS[form] | ts.FLAG_SYNC)
        out.append(cbuf)
    return out


CALLS = [
    # (groups, level, impl, form): two in a row each, the second smaller
    ((dict(m=(40, 6), n=(36, 5), k=(24, 7)), dict(m=(10, 6), n=(9, 5), k=(12, 3))), 1, "abc", "staged"),
    ((dict(m=(48, 4), n=(44, 4), k=(40, 4)), dict(m=(12,), n=(20,), k=(18,))), 2, "abc", "staged_all"),
    ((dict(m=(64, 3), n=(50, 3), k=(33, 5)), dict(m=(16,), n=(12,), k=(40,))), 2, "naive", ""),
    ((dict(m=(70,), n=(66,), k=(80,), batch=(3,)), dict(m=(20,), n=(24,), k=(10,), batch=(2,))), 1, "ab", ""),
    ((dict(m=(32, 32), n=(32, 32), k=(32, 32)), dict(m=(8, 8), n=(8, 8), k=(8, 8))), 1, "abc", "fused"),
    ((dict(m=(90,), n=(70,), k=(50,)), dict(m=(30,), n=(20,), k=(10,))), 0, "abc", ""),
]


@pytest.mark.parametrize("ci", range(len(CALLS)))
def test_noalloc_equals_regular_twice(ci):
    (big, small), level, impl, form = CALLS[ci]
    for j, groups in enumerate((big, small)):
        case = cases.make_case(600 + 10 * ci + j, **groups)
        say(f"noalloc L{level} {impl} {form or 'default'} call {j}: {case['idx_C']} = {case['idx_A']} {case['idx_B']}")
        got, want = both(case, level, impl, form, 40 + j)
        assert torch.equal(got, want), (ci, j)


def test_noalloc_fp32_and_c1_size():
    case = cases.make_case(1024, m=(32, 32), n=(32, 32), k=(32, 32))
    say("noalloc c1 size L1 abc staged, FP64 and FP32")
    got, want = both(case, 1, "abc", "staged", 7)
    assert torch.equal(got, want)
    got, want = both(case, 1, "abc", "", 8, dtype=torch.float32)
    assert torch.equal(got, want)


def test_noalloc_staged_sums_that_do_not_fit_run_fused():
    # 2048^3 at L1: A's 7 staged sums alone are 7 * 1024^2 * 8 bytes = 56 MiB, beyond the block with T's
    case = cases.make_case(2048, m=(64, 32), n=(64, 32), k=(64, 32))
    say("noalloc 2048^3 L1 abc staged -> fused")
    got, _ = both(case, 1, "abc", "staged", 9)
    _, fused = both(case, 1, "abc", "fused", 9)
    assert torch.equal(got, fused)


def test_noalloc_errors():
    # L1 AB stages 7 products of 2048 x 2048 doubles: 224 MiB, beyond the 64 MiB block
    case = cases.make_case(2049, m=(64, 64), n=(64, 64), k=(8, 4))
    n = [int(np.prod(case[f"shape_{x}"])) for x in "ABC"]
    A, _ = dev(case["shape_A"], cases.uniform(n[0], 1))
    B, _ = dev(case["shape_B"], cases.uniform(n[1], 2))
    C, cbuf = dev(case["shape_C"], np.zeros(n[2]))
    say("noalloc errors: AB products beyond the block, low precision, host tensors")
    with pytest.raises(ts.MistraError, match="AB needs"):
        ts.tensor_mult_noalloc(1.0, A, case["idx_A"], B, case["idx_B"], 0.0, C, case["idx_C"], level=1, impl="ab")
    with pytest.raises(ts.MistraError, match="native precision"):
        ts.tensor_mult_noalloc(1.0, A, case["idx_A"], B, case["idx_B"], 0.0, C, case["idx_C"], level=1,
                               precision="bf16x3")
    h = np.zeros((4, 4))
    with pytest.raises(ts.MistraError, match="device tensors"):
        ts.tensor_mult_noalloc(1.0, h, "ik", h, "kj", 0.0, h.copy(), "ij", level=1)
    torch.cuda.synchronize()
    assert torch.count_nonzero(cbuf) == 0
    # the block still works after the refused calls
    small = cases.make_case(2050, m=(20,), n=(30,), k=(40,))
    got, want = both(small, 1, "abc", "", 3)
    assert torch.equal(got, want)
