"""Batched contractions (mistra_batch_mult): grouped launches and HIP-graph replay.

Bars: the batch result equals the individual calls' results bit for bit (the grouped kernel runs the
same launches with the same per-element order); FP32 entries are within 1e-4 (relative Frobenius) of
an FP64 einsum; graph replays, with the program's own eager work and synchronizations between them,
each equal the eager batch.
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


def dev(shape, flat, dtype):
    buf = torch.from_numpy(np.ascontiguousarray(flat)).to(device="cuda", dtype=dtype)
    return buf.as_strided(tuple(shape), tuple(cases.colmajor_strides(shape))), buf


def make_entries(batch, dtype, level, seed):
    ents, bufs = [], []
    for i, case in enumerate(batch):
        n = [int(np.prod(case[f"shape_{x}"])) for x in "ABC"]
        a, b, c = cases.uniform(n[0], seed + 3 * i), cases.uniform(n[1], seed + 3 * i + 1), cases.uniform(n[2], seed + 3 * i + 2)
        A, _ = dev(case["shape_A"], a, dtype)
        B, _ = dev(case["shape_B"], b, dtype)
        C, cbuf = dev(case["shape_C"], c, dtype)
        ents.append(dict(A=A, idx_A=case["idx_A"], B=B, idx_B=case["idx_B"], C=C, idx_C=case["idx_C"], alpha=1.0,
                         beta=0.5, level=level, impl="abc"))
        bufs.append((cbuf, a, b, c))
    return ents, bufs


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("level", [0, 1, 2])
def test_batch_equals_individual_calls(dtype, level):
    batch = cases.c4_batch(11 + level, count=24, lo=16, hi=600)
    say("batch vs individual", dtype, level, len(batch))
    ents, bufs = make_entries(batch, dtype, level, 100)
    ents2, bufs2 = make_entries(batch, dtype, level, 100)
    ts.batch_mult(ents, flags=ts.FLAG_SYNC)
    # batch entries run ABC in its fused form (every launch a leaf launch); the same form called one
    # by one gives the same bits
    for e in ents2:
        ts.tensor_mult(e["alpha"], e["A"], e["idx_A"], e["B"], e["idx_B"], e["beta"], e["C"], e["idx_C"],
                       level=level, impl="abc", flags=ts.FLAG_SYNC | ts.FLAG_FUSED_SUMS)
    for (c1, *_), (c2, *_) in zip(bufs, bufs2):
        assert torch.equal(c1, c2)
    tol = 1e-4 if dtype == torch.float32 else 1e-12
    for case, (cb, a, b, c) in zip(batch, bufs):
        A, B, C = (cases.host_tensor(case[f"shape_{x}"], v) for x, v in zip("ABC", (a, b, c)))
        want = cases.einsum_ref(case, A, B, C, 1.0, 0.5)
        got = cases.host_tensor(case["shape_C"], cb.double().cpu().numpy())
        assert cases.rel_err(got, want) <= tol


def test_graph_replays_match_eager():
    batch = cases.c4_batch(5, count=16, lo=32, hi=700)
    ents, bufs = make_entries(batch, torch.float32, 1, 300)
    ents2, bufs2 = make_entries(batch, torch.float32, 1, 300)
    stream = torch.cuda.Stream()
    b = ts.Batch(ents)
    b2 = ts.Batch(ents2)
    for rep in range(4):
        say("graph replay", rep)
        with torch.cuda.stream(stream):
            b.run(flags=ts.FLAG_GRAPH | ts.FLAG_SYNC, stream=stream)
        b2.run(flags=ts.FLAG_SYNC)  # eager batch
        torch.cuda.synchronize()
        junk = torch.ones(1 << 20, device="cuda") * rep  # the program's own eager work between replays
        torch.cuda.synchronize()
        del junk
        for (c1, *_), (c2, *_) in zip(bufs, bufs2):
            assert torch.equal(c1, c2), rep


def test_batch_with_individual_entries():
    # host data and general index classes run as individual calls inside a batch
    case = cases.make_case(9, m=(12,), n=(10,), k=(14,), a_only=(3,))
    n = [int(np.prod(case[f"shape_{x}"])) for x in "ABC"]
    a, b, c = cases.small_ints(n[0], 1), cases.small_ints(n[1], 2), cases.small_ints(n[2], 3)
    A, _ = dev(case["shape_A"], a, torch.float64)
    B, _ = dev(case["shape_B"], b, torch.float64)
    C, cbuf = dev(case["shape_C"], c, torch.float64)
    plain = cases.make_case(10, m=(40,), n=(30,), k=(50,))
    pn = [int(np.prod(plain[f"shape_{x}"])) for x in "ABC"]
    pa, pb, pc = cases.small_ints(pn[0], 4), cases.small_ints(pn[1], 5), cases.small_ints(pn[2], 6)
    hA, hB = cases.host_tensor(plain["shape_A"], pa.copy()), cases.host_tensor(plain["shape_B"], pb.copy())
    hC = cases.host_tensor(plain["shape_C"], pc.copy())
    ts.batch_mult([dict(A=A, idx_A=case["idx_A"], B=B, idx_B=case["idx_B"], C=C, idx_C=case["idx_C"], beta=1.0),
                   dict(A=hA, idx_A=plain["idx_A"], B=hB, idx_B=plain["idx_B"], C=hC, idx_C=plain["idx_C"],
                        beta=1.0)], flags=ts.FLAG_SYNC)
    want = cases.einsum_ref(case, *(cases.host_tensor(case[f"shape_{x}"], v) for x, v in zip("ABC", (a, b, c))))
    assert np.array_equal(cases.host_tensor(case["shape_C"], cbuf.cpu().numpy()), want)
    want2 = cases.einsum_ref(plain, *(cases.host_tensor(plain[f"shape_{x}"], v) for x, v in zip("ABC", (pa, pb, pc))))
    assert np.array_equal(np.array(hC), want2)
