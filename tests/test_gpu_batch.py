"""Batched contractions (mistra_batch_mult): the persistent launch, the grouped launches and HIP-graph
replay of both.

Bars:
  * persistent launch (default): small-integer entries equal numpy.einsum bit for bit at L0/1/2 (packing
    and the leaf kernel are exact on them); uniform entries within 1e-13 (FP64) / 1e-6 (FP32) of the
    individual calls and within 1e-12 / 1e-4 of an FP64 einsum; repeated calls of one cached batch give
    the same bits (the counters reset themselves); a forced unreachable wait gives up and is reported,
    and the next call is right again;
  * grouped launches (MISTRA_FLAG_BATCH_GROUPED): the batch equals the individual calls of the fused form
    bit for bit (the same launches with the same per-element order);
  * graph replays, with the program's own eager work and synchronizations between them, each equal the
    eager batch.
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
def test_grouped_batch_equals_individual_calls(dtype, level):
    batch = cases.c4_batch(11 + level, count=24, lo=16, hi=600)
    say("grouped batch vs individual", dtype, level, len(batch))
    ents, bufs = make_entries(batch, dtype, level, 100)
    ents2, bufs2 = make_entries(batch, dtype, level, 100)
    ts.batch_mult(ents, flags=ts.FLAG_SYNC | ts.FLAG_BATCH_GROUPED)
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


@pytest.mark.parametrize("form", [0, ts.FLAG_BATCH_GROUPED])
def test_graph_replays_match_eager(form):
    batch = cases.c4_batch(5, count=16, lo=32, hi=700)
    ents, bufs = make_entries(batch, torch.float32, 1, 300)
    ents2, bufs2 = make_entries(batch, torch.float32, 1, 300)
    stream = torch.cuda.Stream()
    b = ts.Batch(ents)
    b2 = ts.Batch(ents2)
    for rep in range(4):
        say("graph replay", form, rep)
        with torch.cuda.stream(stream):
            b.run(flags=ts.FLAG_GRAPH | ts.FLAG_SYNC | form, stream=stream)
        b2.run(flags=ts.FLAG_SYNC | form)  # eager batch
        torch.cuda.synchronize()
        junk = torch.ones(1 << 20, device="cuda") * rep  # the program's own eager work between replays
        torch.cuda.synchronize()
        del junk
        for (c1, *_), (c2, *_) in zip(bufs, bufs2):
            assert torch.equal(c1, c2), rep


@pytest.mark.parametrize("form", [0, ts.FLAG_BATCH_GROUPED])
def test_batch_with_individual_entries(form):
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
                        beta=1.0)], flags=ts.FLAG_SYNC | form)
    want = cases.einsum_ref(case, *(cases.host_tensor(case[f"shape_{x}"], v) for x, v in zip("ABC", (a, b, c))))
    assert np.array_equal(cases.host_tensor(case["shape_C"], cbuf.cpu().numpy()), want)
    want2 = cases.einsum_ref(plain, *(cases.host_tensor(plain[f"shape_{x}"], v) for x, v in zip("ABC", (pa, pb, pc))))
    assert np.array_equal(np.array(hC), want2)


# shapes with fringes at every level (odd lengths), multi-index groups that are not contiguous (C
# sub-blocks that are not translates), K-fast and M-fast operands
PERSIST_CASES = [dict(m=(13, 7), n=(11, 5), k=(9, 6)), dict(m=(64,), n=(48,), k=(40,)),
                 dict(m=(33, 4), n=(70,), k=(3, 25)), dict(m=(128, 2), n=(130,), k=(200,)),
                 dict(m=(5,), n=(260,), k=(7, 9)), dict(m=(96, 3), n=(17, 9), k=(64,))]


def persist_batch(dtype, level, gen, seed, alpha=1.5, beta=0.5):
    batch = [cases.make_case(seed + i, **g) for i, g in enumerate(PERSIST_CASES)]
    batch += cases.c4_batch(seed, count=10, lo=16, hi=500)
    ents, bufs = [], []
    for i, case in enumerate(batch):
        n = [int(np.prod(case[f"shape_{x}"])) for x in "ABC"]
        a, b, c = gen(n[0], seed + 3 * i), gen(n[1], seed + 3 * i + 1), gen(n[2], seed + 3 * i + 2)
        A, _ = dev(case["shape_A"], a, dtype)
        B, _ = dev(case["shape_B"], b, dtype)
        C, cbuf = dev(case["shape_C"], c, dtype)
        ents.append(dict(A=A, idx_A=case["idx_A"], B=B, idx_B=case["idx_B"], C=C, idx_C=case["idx_C"], alpha=alpha,
                         beta=beta, level=level, impl="abc"))
        bufs.append((cbuf, a, b, c))
    return batch, ents, bufs


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("level", [0, 1, 2])
def test_persistent_batch_exact_on_integers(dtype, level):
    batch, ents, bufs = persist_batch(dtype, level, cases.small_ints, 700 + level, alpha=2.0, beta=-1.0)
    say("persistent exact", dtype, level, len(batch))
    ts.batch_mult(ents, flags=ts.FLAG_SYNC)
    for case, (cb, a, b, c) in zip(batch, bufs):
        A, B, C = (cases.host_tensor(case[f"shape_{x}"], v) for x, v in zip("ABC", (a, b, c)))
        want = cases.einsum_ref(case, A, B, C, 2.0, -1.0)
        assert np.array_equal(cases.host_tensor(case["shape_C"], cb.double().cpu().numpy()), want), case


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("level", [0, 1, 2])
def test_persistent_batch_matches_individual_calls(dtype, level):
    batch, ents, bufs = persist_batch(dtype, level, cases.uniform, 800 + level)
    _, ents2, bufs2 = persist_batch(dtype, level, cases.uniform, 800 + level)
    say("persistent vs individual", dtype, level, len(batch))
    b = ts.Batch(ents)
    b.run(flags=ts.FLAG_SYNC)
    for e in ents2:
        ts.tensor_mult(e["alpha"], e["A"], e["idx_A"], e["B"], e["idx_B"], e["beta"], e["C"], e["idx_C"],
                       level=level, impl="abc", flags=ts.FLAG_SYNC)
    near, far = (1e-6, 1e-4) if dtype == torch.float32 else (1e-13, 1e-12)
    first = [cb.clone() for cb, *_ in bufs]
    for case, (cb, a, b_, c), (cb2, *_) in zip(batch, bufs, bufs2):
        assert cases.rel_err(cb.double().cpu().numpy(), cb2.double().cpu().numpy()) <= near, case
        A, B, C = (cases.host_tensor(case[f"shape_{x}"], v) for x, v in zip("ABC", (a, b_, c)))
        want = cases.einsum_ref(case, A, B, C, 1.5, 0.5)
        assert cases.rel_err(cases.host_tensor(case["shape_C"], cb.double().cpu().numpy()), want) <= far, case
    # the same batch again (its tables and counters cached): the same bits, from a C of NaNs with beta = 0
    for e in ents:
        e["beta"] = 0.0
    for rep in range(3):
        for cb, *_ in bufs:
            cb.fill_(float("nan"))
        b.run(flags=ts.FLAG_SYNC)
        for (cb, a, b_, c), case in zip(bufs, batch):
            A, B = (cases.host_tensor(case[f"shape_{x}"], v) for x, v in zip("AB", (a, b_)))
            want = cases.einsum_ref(case, A, B, cases.host_tensor(case["shape_C"], np.zeros_like(c)), 1.5, 0.0)
            assert cases.rel_err(cases.host_tensor(case["shape_C"], cb.double().cpu().numpy()), want) <= far, rep


def test_persistent_batch_reports_waits_that_give_up():
    """Every wait made unreachable from the host (a bias on its target) gives up after a few polls: the
    launch still ends, the synchronous call reports it, and the next call is right."""
    batch, ents, bufs = persist_batch(torch.float64, 1, cases.small_ints, 900, alpha=1.0, beta=0.0)
    ts.debug_set_batch_wait(spin_limit=100, dep_bias=1)
    try:
        with pytest.raises(ts.MistraError, match="gave up"):
            ts.batch_mult(ents, flags=ts.FLAG_SYNC)
    finally:
        ts.debug_set_batch_wait(0, 0)
    ts.batch_mult(ents, flags=ts.FLAG_SYNC)
    for case, (cb, a, b, c) in zip(batch, bufs):
        A, B = (cases.host_tensor(case[f"shape_{x}"], v) for x, v in zip("AB", (a, b)))
        want = cases.einsum_ref(case, A, B, cases.host_tensor(case["shape_C"], np.zeros_like(c)), 1.0, 0.0)
        assert np.array_equal(cases.host_tensor(case["shape_C"], cb.cpu().numpy()), want)


def test_persistent_batch_mixed_types():
    """FP32 and FP64 entries in one batch: the persistent launch takes the first entry's type, the others
    run as individual calls."""
    b32 = [cases.make_case(950 + i, **g) for i, g in enumerate(PERSIST_CASES[:3])]
    b64 = [cases.make_case(960 + i, **g) for i, g in enumerate(PERSIST_CASES[3:])]
    ents, checks = [], []
    for i, (case, dt) in enumerate([(c, torch.float32) for c in b32] + [(c, torch.float64) for c in b64]):
        n = [int(np.prod(case[f"shape_{x}"])) for x in "ABC"]
        a, b, c = cases.small_ints(n[0], 3 * i), cases.small_ints(n[1], 3 * i + 1), cases.small_ints(n[2], 3 * i + 2)
        A, _ = dev(case["shape_A"], a, dt)
        B, _ = dev(case["shape_B"], b, dt)
        C, cbuf = dev(case["shape_C"], c, dt)
        ents.append(dict(A=A, idx_A=case["idx_A"], B=B, idx_B=case["idx_B"], C=C, idx_C=case["idx_C"], beta=1.0,
                         level=1, impl="abc"))
        checks.append((case, cbuf, a, b, c))
    ts.batch_mult(ents, flags=ts.FLAG_SYNC)
    for case, cbuf, a, b, c in checks:
        want = cases.einsum_ref(case, *(cases.host_tensor(case[f"shape_{x}"], v) for x, v in zip("ABC", (a, b, c))))
        assert np.array_equal(cases.host_tensor(case["shape_C"], cbuf.double().cpu().numpy()), want)
