"""GPU: a caller captures one staged contraction into a HIP graph (torch.cuda.graph on its stream).

The staged forms take stream-ordered workspace and, outside capture, run B's sums on a second stream;
under capture they stay on the caller's stream.  Every replay (with the program's own eager work and
synchronizations between replays) must equal the eager call bit for bit.  The plan is created by a
warm-up call before the capture (plan creation uploads tables synchronously).
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


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def dev(shape, flat):
    buf = torch.from_numpy(np.ascontiguousarray(flat)).to(device="cuda", dtype=torch.float64)
    return buf.as_strided(tuple(shape), tuple(cases.colmajor_strides(shape))), buf


@pytest.mark.parametrize("level,impl,flags", [
    (1, "abc", ts.FLAG_OUTER_STAGED), (2, "abc", ts.FLAG_ALL_STAGED), (2, "naive", 0), (1, "ab", 0),
])
def test_graph_replays_of_a_staged_call_match_eager(level, impl, flags):
    case = cases.make_case(900 + level, m=(16, 18), n=(12, 22), k=(20, 13))
    n = [int(np.prod(case[f"shape_{x}"])) for x in "ABC"]
    a, b = cases.uniform(n[0], 51), cases.uniform(n[1], 52)
    A, _ = dev(case["shape_A"], a)
    B, _ = dev(case["shape_B"], b)
    C, cbuf = dev(case["shape_C"], np.zeros(n[2]))
    stream = torch.cuda.Stream()
    say(f"capture: L{level} {impl} flags {flags}")
    with torch.cuda.stream(stream):
        ts.tensor_mult(1.0, A, case["idx_A"], B, case["idx_B"], 0.0, C, case["idx_C"], level=level, impl=impl,
                       flags=flags, stream=stream)
    torch.cuda.synchronize()
    eager = cbuf.clone()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=stream):
        ts.tensor_mult(1.0, A, case["idx_A"], B, case["idx_B"], 0.0, C, case["idx_C"], level=level, impl=impl,
                       flags=flags, stream=stream)
    for rep in range(3):
        say("replay", rep)
        cbuf.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(cbuf, eager), rep
        junk = torch.ones(1 << 20, device="cuda", dtype=torch.float64) * rep  # eager work between replays
        torch.cuda.synchronize()
        del junk
    want = cases.einsum_ref(case, cases.host_tensor(case["shape_A"], a), cases.host_tensor(case["shape_B"], b),
                            cases.host_tensor(case["shape_C"], np.zeros(n[2])), 1.0, 0.0)
    got = cases.host_tensor(case["shape_C"], eager.cpu().numpy())
    assert cases.rel_err(got, want) <= 1e-13
