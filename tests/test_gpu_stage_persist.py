"""GPU: MISTRA_FLAG_STAGE_PERSIST -- every product of a staged ABC plan (level 1 staged, level 2 all
staged) in one persistent launch whose tiles wait for the earlier products that update the same C tile.

  * small integers: bit-exact against numpy.einsum (FP64), with and without fringes;
  * uniform[-1,1): within 1e-13 of the oracle (FP64), within 1e-5 of FP64 einsum (FP32);
  * exactly one timed leaf launch per call (the per-stage form makes one per stage);
  * every C tile's updates follow one fixed product order, so repeated calls agree bit for bit;
  * a wait that gives up (forced from the host: a tiny spin limit and an unreachable target) is reported
    by a synchronous call, and the next call is correct again;
  * graph replays equal the eager call; the allocation-free entry point equals the allocating one.
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

STAGED = {1: ts.FLAG_OUTER_STAGED, 2: ts.FLAG_ALL_STAGED}
# Shapes the staged form takes: K extents whose leaf K (k / 2^L) is a whole number of K-tiles (16 FP64 /
# 32 FP32 elements), and 2^L splits that fall on whole multiples of each group's faster indices (so the
# sub-blocks are translates of each other); (130,) x (66,) leaves fringes at both levels
SHAPES = [
    dict(m=(64,), n=(64,), k=(64,)),
    dict(m=(130,), n=(66,), k=(128,)),
    dict(m=(16, 8), n=(8, 12), k=(16, 16)),
    dict(m=(8, 40), n=(24, 12), k=(8, 32)),
]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def dev(shape, flat, dtype=torch.float64):
    buf = torch.from_numpy(np.ascontiguousarray(flat)).to(device="cuda", dtype=dtype)
    return buf.as_strided(tuple(shape), tuple(cases.colmajor_strides(shape))), buf


def inputs(case, seed, ints):
    n = [int(np.prod(case[f"shape_{x}"])) for x in "ABC"]
    gen = cases.small_ints if ints else cases.uniform
    return gen(n[0], seed), gen(n[1], seed + 1), gen(n[2], seed + 2)


def run(case, a, b, c, alpha, beta, level, flags, dtype=torch.float64, timed=False):
    A, _ = dev(case["shape_A"], a, dtype)
    B, _ = dev(case["shape_B"], b, dtype)
    C, cbuf = dev(case["shape_C"], c, dtype)
    if timed:
        ts.kernel_timing(True)
        ts.kernel_timing_read(reset=True)
    ts.tensor_mult(alpha, A, case["idx_A"], B, case["idx_B"], beta, C, case["idx_C"], level=level, impl="abc",
                   flags=flags | ts.FLAG_SYNC)
    launches = None
    if timed:
        _, launches, _ = ts.kernel_timing_read(reset=True)
        ts.kernel_timing(False)
    return np.array(cases.host_tensor(case["shape_C"], cbuf.double().cpu().numpy())), launches


@pytest.mark.parametrize("level", [1, 2])
@pytest.mark.parametrize("si", range(len(SHAPES)))
def test_integers_exact_and_one_launch(level, si):
    case = cases.make_case(700 + 10 * si + level, **SHAPES[si])
    a, b, c = inputs(case, 40 + si, ints=True)
    say(f"L{level} {SHAPES[si]}: {case['idx_C']} += {case['idx_A']} * {case['idx_B']}")
    ref = cases.einsum_ref(case, cases.host_tensor(case["shape_A"], a), cases.host_tensor(case["shape_B"], b),
                           cases.host_tensor(case["shape_C"], c), 2.0, -1.0)
    got, launches = run(case, a, b, c, 2.0, -1.0, level, STAGED[level] | ts.FLAG_STAGE_PERSIST, timed=True)
    assert np.array_equal(got, ref)
    staged, nstage = run(case, a, b, c, 2.0, -1.0, level, STAGED[level], timed=True)
    assert np.array_equal(staged, ref)
    # one persistent launch over the staged products (the per-stage form: 4 at level 1, 16 at level 2)
    assert launches == 1 and nstage > 1, (launches, nstage)


@pytest.mark.parametrize("level", [1, 2])
def test_uniform_against_oracle_and_repeatable(level):
    case = cases.make_case(760 + level, **SHAPES[2])
    a, b, c = inputs(case, 77, ints=False)
    flags = STAGED[level] | ts.FLAG_STAGE_PERSIST
    got, _ = run(case, a, b, c, 0.5, 0.25, level, flags)
    A, B = cases.host_tensor(case["shape_A"], a), cases.host_tensor(case["shape_B"], b)
    C = cases.host_tensor(case["shape_C"], c.copy())
    oracle.tensor_mult(0.5, A, case["idx_A"], B, case["idx_B"], 0.25, C, case["idx_C"], level=level, impl="abc")
    ref = np.array(C)
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    say(f"L{level}: rel err vs oracle {err:.2e}")
    assert err <= 1e-13
    again, _ = run(case, a, b, c, 0.5, 0.25, level, flags)
    assert np.array_equal(again, got)


@pytest.mark.parametrize("level", [1, 2])
def test_fp32(level):
    case = cases.make_case(780 + level, m=(16, 8), n=(8, 12), k=(32, 8))
    a, b, c = inputs(case, 91, ints=False)
    got, launches = run(case, a, b, c, 1.0, 1.0, level, STAGED[level] | ts.FLAG_STAGE_PERSIST, dtype=torch.float32,
                        timed=True)
    ref = cases.einsum_ref(case, cases.host_tensor(case["shape_A"], a), cases.host_tensor(case["shape_B"], b),
                           cases.host_tensor(case["shape_C"], c), 1.0, 1.0)
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    say(f"FP32 L{level}: rel err vs FP64 einsum {err:.2e}, launches {launches}")
    assert err <= 1e-5 and launches == 1


def test_wait_that_gives_up_is_reported():
    case = cases.make_case(790, **SHAPES[0])
    a, b, c = inputs(case, 5, ints=True)
    flags = ts.FLAG_ALL_STAGED | ts.FLAG_STAGE_PERSIST
    ts.debug_set_batch_wait(spin_limit=16, dep_bias=1)
    try:
        with pytest.raises(ts.MistraError, match="gave up"):
            run(case, a, b, c, 1.0, 0.0, 2, flags)
    finally:
        ts.debug_set_batch_wait(0, 0)
    ref = cases.einsum_ref(case, cases.host_tensor(case["shape_A"], a), cases.host_tensor(case["shape_B"], b),
                           cases.host_tensor(case["shape_C"], c), 1.0, 0.0)
    got, _ = run(case, a, b, c, 1.0, 0.0, 2, flags)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("level", [1, 2])
def test_graph_replays_match_eager(level):
    case = cases.make_case(800 + level, **SHAPES[2])
    n = [int(np.prod(case[f"shape_{x}"])) for x in "ABC"]
    A, _ = dev(case["shape_A"], cases.uniform(n[0], 61))
    B, _ = dev(case["shape_B"], cases.uniform(n[1], 62))
    C, cbuf = dev(case["shape_C"], np.zeros(n[2]))
    flags = STAGED[level] | ts.FLAG_STAGE_PERSIST
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        ts.tensor_mult(1.0, A, case["idx_A"], B, case["idx_B"], 0.0, C, case["idx_C"], level=level, impl="abc",
                       flags=flags, stream=stream)
    torch.cuda.synchronize()
    eager = cbuf.clone()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=stream):
        ts.tensor_mult(1.0, A, case["idx_A"], B, case["idx_B"], 0.0, C, case["idx_C"], level=level, impl="abc",
                       flags=flags, stream=stream)
    for rep in range(3):
        say("replay", rep)
        cbuf.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(cbuf, eager), rep


def test_allocation_free_entry_point_matches():
    case = cases.make_case(820, **SHAPES[3])
    n = [int(np.prod(case[f"shape_{x}"])) for x in "ABC"]
    a, b = cases.uniform(n[0], 63), cases.uniform(n[1], 64)
    flags = ts.FLAG_ALL_STAGED | ts.FLAG_STAGE_PERSIST
    out = []
    for fn in (ts.tensor_mult, ts.tensor_mult_noalloc):
        A, _ = dev(case["shape_A"], a)
        B, _ = dev(case["shape_B"], b)
        C, cbuf = dev(case["shape_C"], np.zeros(n[2]))
        fn(1.0, A, case["idx_A"], B, case["idx_B"], 0.0, C, case["idx_C"], level=2, impl="abc",
           flags=flags | ts.FLAG_SYNC)
        out.append(cbuf.cpu())
    assert torch.equal(out[0], out[1])
