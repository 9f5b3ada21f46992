"""Autotuner on the GPU: a measured choice is one of the candidates, is cached per shape, and calls
with level=-1 / impl="auto" run it and stay within the FP64 tolerance of the classical result."""
import numpy as np
import pytest

import cases
import tblis_strassen as ts

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

from test_gpu_parity import dev  # noqa: E402


@pytest.mark.parametrize("groups", [dict(m=(32, 32), n=(32, 32), k=(32, 32)), dict(m=(24, 40), n=(56,), k=(30, 20))])
def test_measured_choice_runs_and_is_accurate(groups):
    case = cases.make_case(77, **groups)
    n = [int(np.prod(case[f"shape_{x}"])) for x in "ABC"]
    A, _ = dev(case["shape_A"], cases.uniform(n[0], 1))
    B, _ = dev(case["shape_B"], cases.uniform(n[1], 2))
    C, cbuf = dev(case["shape_C"], cases.uniform(n[2], 3))
    before = cbuf.clone()
    level, impl, tile = ts.autotune(A, case["idx_A"], B, case["idx_B"], C, case["idx_C"], measure=True)
    print("measured choice", level, impl, tile, flush=True)
    assert level in (0, 1, 2) and impl in (0, 3, 4) and tile in (0, 1, 2, 3)
    assert torch.equal(cbuf, before)  # tuning ran on a scratch copy
    assert ts.autotune(A, case["idx_A"], B, case["idx_B"], C, case["idx_C"]) == (level, impl, tile)  # cached
    ref, rbuf = dev(case["shape_C"], cases.uniform(n[2], 3))
    ts.tensor_mult(1.0, A, case["idx_A"], B, case["idx_B"], 1.0, ref, case["idx_C"], level=0, flags=ts.FLAG_SYNC)
    ts.tensor_mult(1.0, A, case["idx_A"], B, case["idx_B"], 1.0, C, case["idx_C"], level=-1, impl="auto",
                   flags=ts.FLAG_SYNC)
    err = float(torch.linalg.norm(cbuf - rbuf) / torch.linalg.norm(rbuf))
    assert err <= 5e-15 * 2 ** level


def _timed_rounds(fns, rounds=5, calls=3):
    """Per-call time in ms of each callable: the minimum over `rounds` rounds of `calls` back-to-back calls
    (steady state, as a caller's loop runs them), every callable once per round with the start rotated from
    round to round, after one untimed call of each -- so drift of the clocks over the run and the cost of
    following another candidate fall on all alike."""
    for fn in fns:
        fn()
    torch.cuda.synchronize()
    n = len(fns)
    best = [1e30] * n
    for r in range(rounds):
        for j in range(n):
            i = (j + r * n // rounds) % n
            t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            t0.record()
            for _ in range(calls):
                fns[i]()
            t1.record()
            torch.cuda.synchronize()
            best[i] = min(best[i], t0.elapsed_time(t1) / calls)
    return best


def _c5(which):
    lab = {"fwd": ("dcaf", "dfbe", "abce"), "dA": ("abce", "dfbe", "dcaf"), "dB": ("dcaf", "abce", "dfbe")}[which]
    return dict(idx_A=lab[0], idx_B=lab[1], idx_C=lab[2], shape_A=[64] * 4, shape_B=[64] * 4, shape_C=[64] * 4)


@pytest.mark.slow
@pytest.mark.parametrize("workload", ["c2", "c3", "c5fwd", "c5dA", "c5dB"])
def test_model_choice_within_2pct_of_measured_best(workload):
    """Choices of (level, impl, register tile) at the benchmark configurations against the fastest of every
    candidate timed here on the device (minimum of 5 rounds over all candidates):
      * the measured mode's choice (its own timing: the minimum over 3 rounds of 3 back-to-back calls per
        candidate) runs within 3% of it: two separate timings in one process, whose differences on the
        1 ms c5 calls reach 2%;
      * the analytic model's choice runs within 2% of it at c3 and c5, and within 5% at c2, where the
        register tiles of L2 Naive's dense GEMM trade places from one process to the next (each of the
        three was fastest in some run, the spread up to 4%, while the other candidates repeat within
        0.5%), so that no fixed tile stays within 2% everywhere."""
    import bench

    if workload == "c2":
        case, dtype = bench.make_workload("c2"), torch.float64
    elif workload == "c3":
        case, dtype = cases.make_case(4242, m=(128, 128), n=(128, 128), k=(32, 32)), torch.float64
    else:
        case, dtype = _c5(workload[2:]), torch.float32
    T = {x: bench.device_tensor(case[f"shape_{x}"], dtype, torch) for x in "ABC"}
    for x, (_, buf) in T.items():
        ts.fill_uniform(buf, ord(x))
    (A, _), (B, _), (C, _) = T["A"], T["B"], T["C"]
    cands = ts.autotune_candidates(A, case["idx_A"], B, case["idx_B"], C, case["idx_C"])

    def call(level, impl, tile):
        def fn():
            ts.debug_set_tile(tile)
            ts.tensor_mult(1.0, A, case["idx_A"], B, case["idx_B"], 0.0, C, case["idx_C"], level=level, impl=impl)
        return fn

    try:
        times = _timed_rounds([call(level, impl, tile) for level, impl, tile, _ in cands])
    finally:
        ts.debug_set_tile(-1)
    for (level, impl, tile, model_ms), t in zip(cands, times):
        print(f"{workload} L{level} impl={impl} tile={tile}: model {model_ms:.3f} ms, measured {t:.3f} ms", flush=True)
    best = min(times)
    print(f"{workload}: model choice {times[0]:.3f} ms, measured best {best:.3f} ms ({times[0] / best:.4f})", flush=True)
    try:
        measured = ts.autotune(A, case["idx_A"], B, case["idx_B"], C, case["idx_C"], measure=True)
    finally:
        ts.debug_set_tile(-1)
    mi = [(lv, im, ti) for lv, im, ti, _ in cands].index(measured)
    print(f"{workload}: measured choice L{measured[0]} impl={measured[1]} tile={measured[2]} {times[mi]:.3f} ms "
          f"({times[mi] / best:.4f})", flush=True)
    assert times[mi] <= 1.03 * best
    assert times[0] <= (1.05 if workload == "c2" else 1.02) * best
