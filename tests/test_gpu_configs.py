"""The benchmark configurations at their real sizes, on the GPU.

C1 (1024^3, 2+2+2 indices of 32, seeded permutations): the GPU L=0 result is pinned to the CPU oracle
(within 1e-13), and L=1 ABC (and AB / Naive) agree with it within 1e-12.
C2 (8192^3, 3+3+3 indices {16,16,32} in seeded order): too large for the oracle in a test, so every
level/impl is checked against the GPU L=0 result (tolerance 1e-12 at L=1, 5e-12 at L=2), plus
size-independent properties: linearity in A (C(2A) = 2 C(A) exactly in FP64, a power-of-two scale) and
agreement of the three impls with each other.
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


def workload(groups, seed=12345):
    rng = np.random.default_rng(seed)
    g = [tuple(int(x) for x in rng.permutation(groups)) for _ in range(3)]
    return cases.make_case(seed, m=g[0], n=g[1], k=g[2])


def dev_empty(shape):
    n = int(np.prod(shape))
    buf = torch.empty(n, dtype=torch.float64, device="cuda")
    return buf.as_strided(tuple(shape), tuple(cases.colmajor_strides(shape))), buf


def test_c1_against_oracle():
    case = workload((32, 32))
    m, n, k = cases.mnk(case)
    assert (m, n, k) == (1024, 1024, 1024)
    say(f"[C1] {case['idx_C']} += {case['idx_A']} * {case['idx_B']}")
    a, b = cases.uniform(m * k, 1), cases.uniform(k * n, 2)
    A, abuf = dev_empty(case["shape_A"])
    B, bbuf = dev_empty(case["shape_B"])
    abuf.copy_(torch.from_numpy(a))
    bbuf.copy_(torch.from_numpy(b))
    results = {}
    for level, impl in ((0, "abc"), (1, "abc"), (1, "ab"), (1, "naive"), (2, "abc")):
        C, cbuf = dev_empty(case["shape_C"])
        cbuf.zero_()
        ts.tensor_mult(1.0, A, case["idx_A"], B, case["idx_B"], 1.0, C, case["idx_C"], level=level, impl=impl,
                       flags=ts.FLAG_SYNC)
        results[(level, impl)] = cbuf.cpu().numpy()
    # oracle L=0 and L=1 ABC on the same inputs
    for level in (0, 1):
        Ch = cases.host_tensor(case["shape_C"], np.zeros(m * n))
        oracle.tensor_mult(1.0, cases.host_tensor(case["shape_A"], a), case["idx_A"],
                           cases.host_tensor(case["shape_B"], b), case["idx_B"], 1.0, Ch, case["idx_C"],
                           level=level, impl="abc")
        ref = np.asarray(Ch).ravel(order="F")
        err = cases.rel_err(results[(level, "abc")], ref)
        say(f"[C1] L={level} GPU vs oracle: {err:.2e}")
        assert err <= 1e-13
    c0 = results[(0, "abc")]
    for key, c in results.items():
        err = cases.rel_err(c, c0)
        say(f"[C1] L={key[0]} {key[1]} vs GPU L=0: {err:.2e}")
        assert err <= (1e-12 if key[0] == 1 else 5e-12 if key[0] == 2 else 0.0)


@pytest.mark.slow
def test_c2_full_size():
    case = workload((16, 16, 32))
    m, n, k = cases.mnk(case)
    assert (m, n, k) == (8192, 8192, 8192)
    say(f"[C2] {case['idx_C']} += {case['idx_A']} * {case['idx_B']}")
    A, abuf = dev_empty(case["shape_A"])
    B, bbuf = dev_empty(case["shape_B"])
    ts.fill_uniform(abuf, 1)
    ts.fill_uniform(bbuf, 2)
    C0, c0buf = dev_empty(case["shape_C"])
    c0buf.zero_()
    ts.tensor_mult(1.0, A, case["idx_A"], B, case["idx_B"], 1.0, C0, case["idx_C"], level=0, flags=ts.FLAG_SYNC)
    C, cbuf = dev_empty(case["shape_C"])
    outs = {}
    for level in (1, 2):
        for impl in ("abc", "ab", "naive"):
            cbuf.zero_()
            ts.tensor_mult(1.0, A, case["idx_A"], B, case["idx_B"], 1.0, C, case["idx_C"], level=level, impl=impl,
                           flags=ts.FLAG_SYNC)
            err = ts.rel_error(cbuf, c0buf)
            say(f"[C2] L={level} {impl}: rel err vs L=0 {err:.3e}")
            assert err <= (1e-12 if level == 1 else 5e-12)
            if impl == "abc":
                outs[level] = cbuf.clone()
    # linearity with an exact scale: alpha = 2 doubles every rounding step exactly
    cbuf.zero_()
    ts.tensor_mult(2.0, A, case["idx_A"], B, case["idx_B"], 1.0, C, case["idx_C"], level=1, impl="abc",
                   flags=ts.FLAG_SYNC)
    assert torch.equal(cbuf, 2.0 * outs[1])
