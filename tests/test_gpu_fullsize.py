"""The benchmark configurations at full size on the GPU, each pinned to the oracle through a sampled FP64
check: 4096 seeded C elements recomputed on the host over the oracle's scatter vectors (tests/sampled.py).

  C2  m=n=k=8192, 3 indices {16,16,32} per group in seeded order: L0, and L1 / L2 x ABC (staged,
      all-staged, fused) / AB / Naive, at the stated bars 1e-12 (L1) and 5e-12 (L2); L0 at 1e-13.
  C3  m=n=16384, k=1024 (groups 128x128, 128x128, 32x32), A, B uniform times 2^u, u in [-8, 8]: every
      mode -- FP64 L0/1/2, BF16 hi/lo L1/2, Ozaki BF16 and INT8 L0/1/2, FP8 block-scaled L1.
  C4  the 256-entry FP32 batch (tests/cases.py:c4_batch(2024)): the persistent launch within 1e-5 of the
      loop of individual calls, the grouped launches equal to it bit for bit, and every entry against FP64
      at 1e-4.
  C5  C[abce] = A[dcaf] B[dfbe], all lengths 64, forward + dA + dB with 1-level ABC Strassen in FP32 and
      in split-BF16 (FP64 and FP32 tensors), each against FP64 at 1e-4.
"""
import sys

import numpy as np
import pytest

import cases
import sampled
import tblis_strassen as ts

torch = pytest.importorskip("torch")
pytestmark = [pytest.mark.gpu, pytest.mark.slow]

if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def dev_empty(shape, dtype=torch.float64):
    n = int(np.prod(shape))
    buf = torch.empty(n, dtype=dtype, device="cuda")
    return buf.as_strided(tuple(shape), tuple(cases.colmajor_strides(shape))), buf


def seeded_workload(groups, seed=12345):
    """The benchmark's workload: each group's index lengths in a seeded order, seeded permutations."""
    rng = np.random.default_rng(seed)
    g = [tuple(int(x) for x in rng.permutation(groups)) for _ in range(3)]
    return cases.make_case(seed, m=g[0], n=g[1], k=g[2])


def gather(buf, off):
    return buf[torch.from_numpy(off).to("cuda")].double().cpu().numpy()


def test_c2_sampled_oracle_pin():
    case = seeded_workload((16, 16, 32))
    assert cases.mnk(case) == (8192, 8192, 8192)
    say(f"[C2] {case['idx_C']} += {case['idx_A']} * {case['idx_B']}")
    A, abuf = dev_empty(case["shape_A"])
    B, bbuf = dev_empty(case["shape_B"])
    ts.fill_uniform(abuf, 1)
    ts.fill_uniform(bbuf, 2)
    off, ref = sampled.sample(case, abuf.cpu().numpy(), bbuf.cpu().numpy(), nsamp=4096, seed=2)
    C, cbuf = dev_empty(case["shape_C"])
    forms = {"": 0, "staged": ts.FLAG_OUTER_STAGED, "staged_all": ts.FLAG_ALL_STAGED, "fused": ts.FLAG_FUSED_SUMS}
    variants = [(0, "abc", ""), (1, "abc", "staged"), (1, "abc", "fused"), (1, "ab", ""), (1, "naive", ""),
                (2, "abc", "staged_all"), (2, "abc", "staged"), (2, "abc", "fused"), (2, "ab", ""), (2, "naive", "")]
    for level, impl, form in variants:
        cbuf.fill_(float("nan"))  # beta = 0 must overwrite
        ts.tensor_mult(1.0, A, case["idx_A"], B, case["idx_B"], 0.0, C, case["idx_C"], level=level, impl=impl,
                       flags=forms[form] | ts.FLAG_SYNC)
        err = sampled.rel_err(gather(cbuf, off), ref)
        say(f"[C2] L={level} {impl} {form or 'default'}: sampled rel err vs FP64 {err:.3e}")
        assert err <= {0: 1e-13, 1: 1e-12, 2: 5e-12}[level]


C3_MODES = [("native", 0, 1e-12), ("native", 1, 1e-12), ("native", 2, 5e-12), ("bf16x3", 1, 5e-4),
            ("bf16x3", 2, 5e-4), ("ozaki_bf16", 0, 1e-10), ("ozaki_bf16", 1, 1e-10), ("ozaki_bf16", 2, 1e-10),
            ("ozaki_int8", 0, 1e-10), ("ozaki_int8", 1, 1e-10), ("ozaki_int8", 2, 1e-10), ("fp8_block", 1, 0.1)]


def test_c3_every_precision_mode_full_size():
    case = cases.make_case(4242, m=(128, 128), n=(128, 128), k=(32, 32))
    assert cases.mnk(case) == (16384, 16384, 1024)
    say(f"[C3] {case['idx_C']} += {case['idx_A']} * {case['idx_B']}")
    A, abuf = dev_empty(case["shape_A"])
    B, bbuf = dev_empty(case["shape_B"])
    ts.fill_uniform(abuf, 1, emax=8)
    ts.fill_uniform(bbuf, 2, emax=8)
    off, ref = sampled.sample(case, abuf.cpu().numpy(), bbuf.cpu().numpy(), nsamp=4096, seed=3)
    C, cbuf = dev_empty(case["shape_C"])
    for prec, level, bar in C3_MODES:
        cbuf.fill_(float("nan"))
        ts.tensor_mult(1.0, A, case["idx_A"], B, case["idx_B"], 0.0, C, case["idx_C"], level=level, impl="abc",
                       precision=prec, flags=ts.FLAG_SYNC)
        err = sampled.rel_err(gather(cbuf, off), ref)
        say(f"[C3] {prec} L={level}: sampled rel err vs FP64 {err:.3e} (bar {bar:.0e})")
        assert err <= bar


def test_c4_full_batch():
    batch = cases.c4_batch(2024)
    assert len(batch) == 256
    say("[C4] 256 FP32 contractions")
    ents, data = [], []
    for i, c in enumerate(batch):
        A, ab = dev_empty(c["shape_A"], torch.float32)
        B, bb = dev_empty(c["shape_B"], torch.float32)
        C, cb = dev_empty(c["shape_C"], torch.float32)
        ts.fill_uniform(ab, 3 * i + 1)
        ts.fill_uniform(bb, 3 * i + 2)
        cb.fill_(float("nan"))
        ents.append(dict(A=A, idx_A=c["idx_A"], B=B, idx_B=c["idx_B"], C=C, idx_C=c["idx_C"], beta=0.0, level=-1,
                         impl="abc"))
        data.append((ab, bb, cb, c))
    b = ts.Batch(ents)
    b.run(ts.FLAG_SYNC)  # one persistent launch
    batch_out = [cb.clone() for _, _, cb, _ in data]
    for _, _, cb, _ in data:
        cb.fill_(float("nan"))
    b.run(ts.FLAG_SYNC | ts.FLAG_BATCH_GROUPED)  # grouped launches
    grouped_out = [cb.clone() for _, _, cb, _ in data]
    # the loop of individual calls (same level and impl choice, the fused form the grouped launches run)
    for e, (_, _, cb, _) in zip(ents, data):
        cb.fill_(float("nan"))
        ts.tensor_mult(1.0, e["A"], e["idx_A"], e["B"], e["idx_B"], 0.0, e["C"], e["idx_C"], level=-1, impl="abc",
                       flags=ts.FLAG_FUSED_SUMS)
    torch.cuda.synchronize()
    worst = 0.0
    for i, ((ab, bb, cb, c), out, gout) in enumerate(zip(data, batch_out, grouped_out)):
        assert torch.equal(gout, cb), f"entry {i}: grouped batch differs from the individual call"
        assert float(torch.linalg.norm(out.double() - cb.double()) / torch.linalg.norm(cb.double())) <= 1e-5, i
        off, ref = sampled.sample(c, ab.double().cpu().numpy(), bb.double().cpu().numpy(), nsamp=256, seed=i)
        err = sampled.rel_err(gather(out, off), ref)
        worst = max(worst, err)
        assert err <= 1e-4, (i, err)
    say(f"[C4] persistent batch within 1e-5 of the loop, grouped batch equal to it; worst sampled rel err vs "
        f"FP64 {worst:.2e}")


@pytest.mark.parametrize("dtype,prec", [(torch.float32, "native"), (torch.float64, "bf16x3"),
                                        (torch.float32, "bf16x3")])
def test_c5_forward_and_adjoints(dtype, prec):
    L = 64
    shape = [L] * 4
    say(f"[C5] {dtype} {prec} L1 ABC forward + dA + dB")
    A, ab = dev_empty(shape, dtype)
    B, bb = dev_empty(shape, dtype)
    dC, db = dev_empty(shape, dtype)
    C, cb = dev_empty(shape, dtype)
    dA, dab = dev_empty(shape, dtype)
    dB, dbb = dev_empty(shape, dtype)
    ts.fill_uniform(ab, 1)
    ts.fill_uniform(bb, 2)
    ts.fill_uniform(db, 3)
    for x in (cb, dab, dbb):
        x.fill_(float("nan"))
    ts.tensor_mult(1.0, A, "dcaf", B, "dfbe", 0.0, C, "abce", level=1, impl="abc", precision=prec)
    ts.tensor_mult_backward(1.0, A, "dcaf", B, "dfbe", dC, "abce", dA=dA, dB=dB, beta=0.0, level=1, impl="abc",
                            precision=prec)
    torch.cuda.synchronize()
    a, b, g = (x.double().cpu().numpy() for x in (ab, bb, db))
    fwd = dict(idx_A="dcaf", idx_B="dfbe", idx_C="abce", shape_A=shape, shape_B=shape, shape_C=shape,
               lengths={c: L for c in "abcdef"})
    gA = dict(idx_A="abce", idx_B="dfbe", idx_C="dcaf", shape_A=shape, shape_B=shape, shape_C=shape,
              lengths=fwd["lengths"])
    gB = dict(idx_A="dcaf", idx_B="abce", idx_C="dfbe", shape_A=shape, shape_B=shape, shape_C=shape,
              lengths=fwd["lengths"])
    for name, case, x, y, out in (("C", fwd, a, b, cb), ("dA", gA, g, b, dab), ("dB", gB, a, g, dbb)):
        off, ref = sampled.sample(case, x, y, nsamp=4096, seed=7)
        err = sampled.rel_err(gather(out, off), ref)
        say(f"[C5] {name}: sampled rel err vs FP64 {err:.2e}")
        assert err <= 1e-4
