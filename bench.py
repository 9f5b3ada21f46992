#!/usr/bin/env python
"""Benchmark: effective TFLOP/s (2mnk/t) of Strassen tensor contraction (L=1, L=2) against the classical
(L=0) contraction on the same MI355X, FP64, at the stated relative error.

Workload (one GPU): m = n = k = 8192; each of the m, n, k groups has 3 indices of lengths {16,16,32} in
seeded order; each tensor's index order is a seeded permutation; A, B ~ uniform[-1,1) (splitmix64,
generated in HBM); C starts at zero and every step is C += A.B (beta = 1, as the reference's
test_strassen does).  A step is one contraction call; inputs are resident in HBM before timing.

Variants: L0 (classical MFMA TC), L1/L2 x {ABC, AB, Naive}; ABC in each of its forms: staged (the outer
level's 7 operand sums of A and of B written to HBM, the inner sums -- none at L1, 2 terms at L2 -- formed
in the leaf kernel), staged_all (at L2 all 49 sums written to HBM) and fused (every sum formed in the
leaf kernel, no workspace); +atomic: every product in one launch, adding into C with FP64 atomic adds
(run-to-run differences in the last bits).  C is updated in the leaf kernel's epilogue in all of them.  Errors are ||C - C_L0||_F / ||C_L0||_F of a
single call from C = 0.  The reported `value` is the fastest Strassen variant (L1 or L2, any impl) that
meets its tolerance (1e-12 at L1, 5e-12 at L2); every variant, ABC included, is listed under `variants`.

roofline: FP64 MFMA bound of the leaf kernel of the reported variant: achieved = leaf flops per launch /
average launch duration (HIP events recorded by the library on the launch stream); traffic comes from the
committed counter summary profiles/traffic_<workload>.json when present.
precision (default run): the split-precision target on the same contraction -- FP64 classical against
Ozaki BF16 / INT8 digit slicing at L1 and L2 and the 3-product BF16 split, each with its error against
the FP64 classical result and its leaf kernel's fraction of that dtype's MFMA peak.
configs (default run): the c1, c3, c4 and c5 lines below, each without the fields this line carries.
cpu_baseline: the oracle's blocked CPU port of the reference's ABC loop nest on the measured
configuration itself (one 8192^3 L1 call, minimum of 3 with a cache flush between), on the host's
physical cores up to the 16 CPUs this process may use, with lscpu's socket and core counts.

Other workloads (one JSON line each, same fields; --workload NAME):
  c1  the same measurement at m=n=k=1024 (2 indices of 32 per group);
  c3  precision modes at m=n=16384, k=1024 (groups 128x128, 128x128, 32x32; A, B uniform times 2^u,
      u in [-8, 8]): FP64 classical and Strassen, BF16 hi/lo split, Ozaki BF16 and INT8 slicing and FP8
      block scaling, each with its error against the FP64 classical result; value = the fastest mode
      within 1e-10;
  c2p the precision modes on the c2 contraction (m=n=k=8192, uniform[-1,1)): FP64 classical and
      Strassen, BF16 hi/lo, Ozaki BF16 and INT8 slicing; value = the fastest mode within 1e-10;
  c4  256 seeded FP32 contractions of uneven shape (tests/cases.py:c4_batch): a loop of individual
      calls, the grouped batch, and its HIP-graph replay; value = the batch throughput;
  c5  C[abce] = A[dcaf] B[dfbe], all lengths 64, FP32: forward + dA + dB with 1-level ABC (and the
      BF16 split on FP64 copies); value = forward+backward TFLOP/s.
"""
import argparse
import gc
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "tblis-strassen_amd"), os.path.join(ROOT, "tests"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

PEAK_FP64_MFMA = 78.6  # TFLOP/s, MI355X published dense FP64 matrix peak
TOL = {0: 0.0, 1: 1e-12, 2: 5e-12}

WORKLOADS = {
    # name: (per-group index lengths, seed)
    "c2": ((16, 16, 32), 8192),
    "c1": ((32, 32), 1024),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(name, seed=12345):
    import numpy as np

    import cases

    lens, _ = WORKLOADS[name]
    rng = np.random.default_rng(seed)
    groups = [tuple(int(x) for x in rng.permutation(lens)) for _ in range(3)]
    return cases.make_case(seed, m=groups[0], n=groups[1], k=groups[2])


def device_tensor(shape, dtype, torch):
    import cases

    n = 1
    for s in shape:
        n *= s
    buf = torch.empty(n, dtype=dtype, device="cuda")
    return buf.as_strided(tuple(shape), tuple(cases.colmajor_strides(shape))), buf


class no_gc:
    """Python's garbage collector held off over a timed region (collected just before it): a full collection
    of this interpreter's heap takes about 60 ms, and one that fell inside the timed steps of a 14 ms call
    once added 4 ms to each of its 10 steps."""

    def __enter__(self):
        gc.collect()
        gc.disable()

    def __exit__(self, *exc):
        gc.enable()


def run_variant(ts, torch, case, A, B, C, level, impl, steps, warmup, stream, flags=0):
    """Time `steps` calls after `warmup` calls; returns (ms/step, kernel ms/launch, launches, leaf flops/launch)."""
    for _ in range(warmup):
        ts.tensor_mult(1.0, A, case["idx_A"], B, case["idx_B"], 1.0, C, case["idx_C"], level=level, impl=impl,
                       stream=stream, flags=flags)
    torch.cuda.synchronize()
    ts.kernel_timing(True)
    ts.kernel_timing_read(reset=True)
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    with no_gc():
        torch.cuda.synchronize()
        start.record(stream)
        for _ in range(steps):
            ts.tensor_mult(1.0, A, case["idx_A"], B, case["idx_B"], 1.0, C, case["idx_C"], level=level, impl=impl,
                           stream=stream, flags=flags)
        stop.record(stream)
        torch.cuda.synchronize()
    ms = start.elapsed_time(stop) / steps
    kms, launches, kflops = ts.kernel_timing_read(reset=True)
    ts.kernel_timing(False)
    return ms, kms / max(1, launches), launches, kflops / max(1, launches)


def host_cpu():
    """lscpu's model, sockets, cores per socket and threads per core, and the CPUs this process may use."""
    info = {}
    try:
        import subprocess

        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            key, _, val = line.partition(":")
            info[key.strip()] = val.strip()
    except Exception:  # noqa: BLE001 - lscpu is informational only
        pass
    sockets = int(info.get("Socket(s)", "0") or 0)
    cores = int(info.get("Core(s) per socket", "0") or 0)
    return {"model": info.get("Model name", "unknown"), "sockets": sockets, "cores_per_socket": cores,
            "threads_per_core": int(info.get("Thread(s) per core", "0") or 0),
            "physical_cores": sockets * cores, "usable_cpus": len(os.sched_getaffinity(0))}


def cpu_baseline(case, threads, repeats=3, level=1):
    """The oracle's blocked 1-level ABC Strassen (a CPU port of the reference's loop nest: packed
    operand-sum panels, register micro-kernel, multi-C update; oracle_tensor_mult_blocked) on the measured
    configuration itself: one call = the whole contraction, timed as the reference's harness does
    (test/test_strassen.cxx:74-95: the minimum over `repeats` runs, caches flushed between runs by writing
    a 256 MiB buffer)."""
    import numpy as np

    import cases
    import oracle

    oracle.load()
    m, n, k = cases.mnk(case)
    a, b = oracle.fill_uniform(m * k, 1), oracle.fill_uniform(k * n, 2)
    c = np.zeros(m * n)
    A, B, C = (cases.host_tensor(case[f"shape_{x}"], v) for x, v in zip("ABC", (a, b, c)))
    trash = np.empty(32 << 20)
    best = float("inf")
    for _ in range(repeats):
        trash.fill(1.0)
        t0 = time.perf_counter()
        oracle.tensor_mult_blocked(1.0, A, case["idx_A"], B, case["idx_B"], 1.0, C, case["idx_C"], level=level,
                                   threads=threads)
        best = min(best, time.perf_counter() - t0)
        log(f"cpu baseline: one L{level} ABC call {time.perf_counter() - t0:.2f} s")
    hw = host_cpu()
    return {
        "value": round(2.0 * m * n * k / best / 1e12, 5),
        "unit": "TFLOP/s",
        "cores": threads,
        "kind": "port",
        "sample": f"the measured workload itself: one {level}-level ABC Strassen call of the oracle's blocked CPU "
                  f"port of the reference loop nest, FP64, m=n=k={m} with the same index layout, minimum of {repeats} "
                  f"calls with a 256 MiB cache flush between them, {threads} OpenMP threads; host "
                  f"{hw['model']}, {hw['sockets']} socket(s) x {hw['cores_per_socket']} cores "
                  f"({hw['physical_cores']} physical), {hw['usable_cpus']} CPUs usable by this process",
        "seconds_per_call": round(best, 3),
        "host": hw,
    }


def timed(torch, fn, steps, warmup, stream):
    for _ in range(warmup):
        fn()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    with no_gc():
        torch.cuda.synchronize()
        start.record(stream)
        for _ in range(steps):
            fn()
        stop.record(stream)
        torch.cuda.synchronize()
    return start.elapsed_time(stop) / steps


def common(args, metric, value, unit, ms, dtype, config, roofline, extra=None):
    out = {
        "metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": dtype, "data": "synthetic (splitmix64 uniform[-1,1), seeded index permutations)", "config": config,
        "roofline": roofline, "cpu_baseline": None,
    }
    if extra:
        out.update(extra)
    return out


PEAK = {"f64": 78.6, "f32": 157.3, "bf16": 2500.0, "fp8": 5000.0, "int8": 5000.0}


def bench_c3(ts, torch, args, stream, workload="c3"):
    """Precision modes; errors vs FP64 classical.  c3: m=n=16384, k=1024 with 2^u-scaled inputs; c2p: the
    c2 contraction (m=n=k=8192, uniform[-1,1) inputs), where the split-precision Strassen target is set."""
    import cases

    if workload == "c3":
        case = cases.make_case(4242, m=(128, 128), n=(128, 128), k=(32, 32))
        emax = 8
        modes = [("fp64", 0), ("fp64", 1), ("fp64", 2), ("bf16x3", 1), ("bf16x3", 2), ("ozaki_bf16", 0),
                 ("ozaki_bf16", 1), ("ozaki_bf16", 2), ("ozaki_int8", 0), ("ozaki_int8", 1), ("ozaki_int8", 2),
                 ("fp8_block", 1)]
    else:
        case = make_workload("c2")
        emax = 0
        modes = [("fp64", 0), ("fp64", 1), ("fp64", 2), ("bf16x3", 1), ("ozaki_bf16", 0), ("ozaki_bf16", 1),
                 ("ozaki_bf16", 2), ("ozaki_int8", 0), ("ozaki_int8", 1), ("ozaki_int8", 2)]
    if args.modes:
        modes = [(x.split(":")[0], int(x.split(":")[1])) for x in args.modes.split(",")]
    m, n, k = cases.mnk(case)
    eff = 2.0 * m * n * k
    A, abuf = device_tensor(case["shape_A"], torch.float64, torch)
    B, bbuf = device_tensor(case["shape_B"], torch.float64, torch)
    ts.fill_uniform(abuf, 1, emax=emax, stream=stream)
    ts.fill_uniform(bbuf, 2, emax=emax, stream=stream)
    Cref, crefbuf = device_tensor(case["shape_C"], torch.float64, torch)
    crefbuf.zero_()
    ts.tensor_mult(1.0, A, case["idx_A"], B, case["idx_B"], 0.0, Cref, case["idx_C"], level=0, stream=stream)
    C, cbuf = device_tensor(case["shape_C"], torch.float64, torch)
    log(f"workload {workload}: m={m} n={n} k={k}")
    res = {}

    def measure(prec, level):
        pr = "native" if prec == "fp64" else prec

        def call():
            ts.tensor_mult(1.0, A, case["idx_A"], B, case["idx_B"], 0.0, C, case["idx_C"], level=level, impl="abc",
                           precision=pr, stream=stream)

        call()
        err = ts.rel_error(cbuf, crefbuf, stream=stream)
        ts.kernel_timing(True)
        ts.kernel_timing_read(reset=True)
        ms = timed(torch, call, args.steps, args.warmup, stream)
        kms, launches, kflops = ts.kernel_timing_read(reset=True)
        ts.kernel_timing(False)
        name = f"{prec}_L{level}"
        peak = {"fp64": PEAK["f64"], "fp8_block": PEAK["fp8"], "ozaki_int8": PEAK["int8"]}.get(prec, PEAK["bf16"])
        leaf = kflops / kms / 1e9 if kms else None
        res[name] = {"precision": prec, "level": level, "ms_per_step": round(ms, 4),
                     "eff_tflops": round(eff / ms / 1e9, 3), "rel_err_vs_fp64": err,
                     "leaf_kernel_tflops": round(leaf, 2) if leaf else None,
                     "leaf_peak_tflops": peak, "leaf_frac": round(leaf / peak, 4) if leaf else None}
        log(f"{name:14s} {ms:8.3f} ms {eff / ms / 1e9:8.2f} eff TFLOP/s  err {err:.2e}  leaf {leaf or 0:.1f} TFLOP/s")

    for prec, level in modes:
        measure(prec, level)
    ok = [v for v in res.values() if v["rel_err_vs_fp64"] <= 1e-10 and v["precision"] != "fp64"]
    if "fp64_L0" not in res:  # a --modes subset: time the classical reference too
        measure("fp64", 0)
    head = max(ok, key=lambda v: v["eff_tflops"]) if ok else res["fp64_L0"]
    hname = f"{head['precision']}_L{head['level']}"
    desc = ("c3: m=n=16384, k=1024, A,B uniform*2^u (u in [-8,8])" if workload == "c3" else
            "c2p: the c2 contraction, m=n=k=8192, A,B uniform[-1,1)")
    dtype = "bf16" if "bf16" in hname else ("int8" if "int8" in hname else "f64")
    return common(args, "effective TFLOP/s (2mnk/t) at <=1e-10 rel-err vs FP64, low-precision leaves",
                  head["eff_tflops"], "TFLOP/s", head["ms_per_step"], dtype,
                  {"workload": f"{desc}; reported mode {hname}",
                   "m": m, "n": n, "k": k, "reported_mode": hname,
                   "fp64_classical_eff_tflops": res["fp64_L0"]["eff_tflops"],
                   "speedup_vs_fp64_classical": round(head["eff_tflops"] / res["fp64_L0"]["eff_tflops"], 4)},
                  {"bound": "mfma", "achieved": head["leaf_kernel_tflops"], "peak": head["leaf_peak_tflops"],
                   "unit": "TFLOP/s", "frac": head["leaf_frac"], "traffic": None},
                  {"modes": res})


def bench_c4(ts, torch, args, stream):
    """256 FP32 contractions: individual-call loop vs grouped batch vs HIP-graph replay."""
    import numpy as np

    import cases

    batch = cases.c4_batch(2024)
    flops = sum(2.0 * float(np.prod([c["lengths"][x] for x in set(c["idx_A"]) | set(c["idx_B"])])) for c in batch)
    ents, refs = [], []
    for i, c in enumerate(batch):
        A, ab = device_tensor(c["shape_A"], torch.float32, torch)
        B, bb = device_tensor(c["shape_B"], torch.float32, torch)
        C, cb = device_tensor(c["shape_C"], torch.float32, torch)
        ts.fill_uniform(ab, 3 * i + 1, stream=stream)
        ts.fill_uniform(bb, 3 * i + 2, stream=stream)
        cb.zero_()
        ents.append(dict(A=A, idx_A=c["idx_A"], B=B, idx_B=c["idx_B"], C=C, idx_C=c["idx_C"], beta=0.0, level=-1,
                         impl="abc"))
        refs.append((ab, bb, cb, c))
    log(f"workload c4: {len(batch)} FP32 contractions, {flops / 1e9:.1f} GFLOP")
    cstream = torch.cuda.Stream()
    torch.cuda.set_stream(cstream)
    stream = cstream

    def loop():
        for e in ents:
            ts.tensor_mult(1.0, e["A"], e["idx_A"], e["B"], e["idx_B"], 0.0, e["C"], e["idx_C"], level=-1,
                           impl="abc", stream=stream)

    b = ts.Batch(ents)
    ms_loop = timed(torch, loop, args.steps, args.warmup, stream)
    ms_grouped = timed(torch, lambda: b.run(ts.FLAG_BATCH_GROUPED, stream), args.steps, args.warmup, stream)
    b.run(ts.FLAG_SYNC, stream)  # builds (and checks) the persistent batch before it is timed
    ts.kernel_timing(True)
    ts.kernel_timing_read(reset=True)
    ms_batch = timed(torch, lambda: b.run(0, stream), args.steps, args.warmup, stream)
    kms, launches, kflops = ts.kernel_timing_read(reset=True)
    ts.kernel_timing(False)
    ms_graph = timed(torch, lambda: b.run(ts.FLAG_GRAPH, stream), args.steps, args.warmup, stream)
    # accuracy: every entry against its FP64 classical result
    worst = 0.0
    for ab, bb, cb, c in refs:
        A64, a64 = device_tensor(c["shape_A"], torch.float64, torch)
        B64, b64 = device_tensor(c["shape_B"], torch.float64, torch)
        C64, c64 = device_tensor(c["shape_C"], torch.float64, torch)
        a64.copy_(ab)
        b64.copy_(bb)
        c64.zero_()
        ts.tensor_mult(1.0, A64, c["idx_A"], B64, c["idx_B"], 0.0, C64, c["idx_C"], level=0, stream=stream)
        torch.cuda.synchronize()
        worst = max(worst, float(torch.linalg.norm(cb.double() - c64) / torch.linalg.norm(c64)))
    rates = {"loop": flops / ms_loop / 1e9, "grouped": flops / ms_grouped / 1e9, "persistent": flops / ms_batch / 1e9,
             "persistent_graph": flops / ms_graph / 1e9}
    log(f"loop {ms_loop:.3f} ms  grouped {ms_grouped:.3f} ms  persistent {ms_batch:.3f} ms  graph {ms_graph:.3f} ms  "
        f"(persistent kernel {kms / max(1, launches):.3f} ms x {launches / max(1, args.steps + args.warmup):.0f} "
        f"per call)  worst err {worst:.2e}")
    best = min(ms_batch, ms_graph)
    torch.cuda.set_stream(torch.cuda.default_stream())
    klaunch = kms / max(1, launches)
    return common(args, "batch throughput TFLOP/s (sum 2mnk / t), 256 FP32 contractions in one persistent launch",
                  round(flops / best / 1e9, 3), "TFLOP/s", best, "f32",
                  {"workload": "c4: 256 seeded FP32 contractions, 3-6 indices, lengths 8-72, group sizes "
                               "log-uniform [64,4096] clamped to the achievable range", "gflop": round(flops / 1e9, 3),
                   "loop_ms": round(ms_loop, 4), "grouped_ms": round(ms_grouped, 4), "batch_ms": round(ms_batch, 4),
                   "graph_ms": round(ms_graph, 4), "persistent_kernel_ms": round(klaunch, 4),
                   "speedup_vs_loop": round(ms_loop / best, 3), "worst_rel_err_vs_fp64": worst,
                   "tflops": {k: round(v, 3) for k, v in rates.items()}},
                  {"bound": "mfma", "achieved": round(kflops / max(1, launches) / klaunch / 1e9, 3) if klaunch else None,
                   "peak": PEAK["f32"], "unit": "TFLOP/s",
                   "frac": round(kflops / max(1, launches) / klaunch / 1e9 / PEAK["f32"], 4) if klaunch else None,
                   "traffic": None})


def bench_c5(ts, torch, args, stream):
    """Forward + adjoints of C[abce] = A[dcaf] B[dfbe] (all lengths 64), FP32 L1 ABC and BF16 split."""
    L = 64
    shape = [L] * 4
    flops = 3 * 2.0 * float(L) ** 6
    res = {}
    ref = None
    for name, dt, level, prec in (("fp64_L0", torch.float64, 0, "native"), ("fp32_L0", torch.float32, 0, "native"),
                                  ("fp32_L1", torch.float32, 1, "native"), ("bf16x3_L1", torch.float64, 1, "bf16x3")):
        A, ab = device_tensor(shape, dt, torch)
        B, bb = device_tensor(shape, dt, torch)
        dC, db = device_tensor(shape, dt, torch)
        C, cb = device_tensor(shape, dt, torch)
        dA, dab = device_tensor(shape, dt, torch)
        dB, dbb = device_tensor(shape, dt, torch)
        ts.fill_uniform(ab, 1, stream=stream)
        ts.fill_uniform(bb, 2, stream=stream)
        ts.fill_uniform(db, 3, stream=stream)

        def step():
            ts.tensor_mult(1.0, A, "dcaf", B, "dfbe", 0.0, C, "abce", level=level, impl="abc", precision=prec,
                           stream=stream)
            ts.tensor_mult_backward(1.0, A, "dcaf", B, "dfbe", dC, "abce", dA=dA, dB=dB, beta=0.0, level=level,
                                    impl="abc", precision=prec, stream=stream)

        ms = timed(torch, step, args.steps, args.warmup, stream)
        outs = (cb.double(), dab.double(), dbb.double())
        if ref is None:
            ref = outs
            errs = {"C": 0.0, "dA": 0.0, "dB": 0.0}
        else:
            errs = {nm: float(torch.linalg.norm(x - y) / torch.linalg.norm(y)) for nm, x, y in zip(("C", "dA", "dB"), outs, ref)}
        res[name] = {"ms_per_step": round(ms, 4), "tflops": round(flops / ms / 1e9, 3), "rel_err_vs_fp64": errs}
        log(f"{name:10s} {ms:8.3f} ms  {flops / ms / 1e9:7.2f} TFLOP/s  errs {errs}")
        del A, B, dC, C, dA, dB, ab, bb, db, cb, dab, dbb
    head = res["fp32_L1"]
    return common(args, "forward+backward TFLOP/s (3 x 2mnk / t), C[abce]=A[dcaf]B[dfbe], lengths 64",
                  head["tflops"], "TFLOP/s", head["ms_per_step"], "f32",
                  {"workload": "c5: FP32 64^6 forward + dA + dB, 1-level ABC", "m": 4096, "n": 4096, "k": 4096,
                   "modes": res},
                  {"bound": "mfma", "achieved": head["tflops"], "peak": PEAK["f32"], "unit": "TFLOP/s",
                   "frac": round(head["tflops"] / PEAK["f32"], 4), "traffic": None})


def bench_tc(ts, torch, args, stream, workload, variant_spec):
    """FP64 Strassen TC at levels 0-2 x ABC / AB / Naive on a c1 / c2 workload; the JSON line of the run."""
    case = make_workload(workload)
    import cases

    m, n, k = cases.mnk(case)
    eff_flops = 2.0 * m * n * k
    log(f"workload {workload}: C[{case['idx_C']}] += A[{case['idx_A']}] B[{case['idx_B']}]  m=n=k={m}")

    A, abuf = device_tensor(case["shape_A"], torch.float64, torch)
    B, bbuf = device_tensor(case["shape_B"], torch.float64, torch)
    ts.fill_uniform(abuf, 1, stream=stream)
    ts.fill_uniform(bbuf, 2, stream=stream)
    Cref, crefbuf = device_tensor(case["shape_C"], torch.float64, torch)
    crefbuf.zero_()
    ts.tensor_mult(1.0, A, case["idx_A"], B, case["idx_B"], 1.0, Cref, case["idx_C"], level=0, stream=stream)
    C, cbuf = device_tensor(case["shape_C"], torch.float64, torch)

    variants = {}
    # a form is one or more of these joined by "+" ("persist": every stage product in one persistent
    # launch instead of one leaf launch per stage)
    forms = {"": 0, "fused": ts.FLAG_FUSED_SUMS, "staged": ts.FLAG_OUTER_STAGED, "staged_all": ts.FLAG_ALL_STAGED,
             "persist": ts.FLAG_STAGE_PERSIST, "atomic": ts.FLAG_ATOMIC_C}
    tiles = {0: "128x128", 1: "64x64", 2: "128x64", 3: "64x128"}
    impls = {0: "abc", 3: "naive", 4: "ab"}
    for spec in variant_spec.split(","):
        choice = None
        if spec == "auto":
            # the autotuner's choice of level, impl and register tile (its analytic model), run as a caller
            # asking for it would: level -1, impl "auto"
            lv_c, im_c, ti_c = ts.autotune(A, case["idx_A"], B, case["idx_B"], C, case["idx_C"])
            choice = {"level": lv_c, "impl": impls.get(im_c, im_c), "tile": tiles.get(ti_c, ti_c)}
            lv, impl, form = "-1", "auto", ""
        else:
            lv, impl, *form = spec.split(":")
            form = form[0] if form else ""
        flags = 0
        for f in form.split("+"):
            flags |= forms[f]
        level = int(lv)
        name = "auto" if choice else f"L{level}_{impl}" + (f"_{form}" if form else "")
        cbuf.zero_()
        ts.tensor_mult(1.0, A, case["idx_A"], B, case["idx_B"], 1.0, C, case["idx_C"], level=level, impl=impl,
                       stream=stream, flags=flags)
        err = ts.rel_error(cbuf, crefbuf, stream=stream)
        cbuf.zero_()
        ms, kms, launches, kflops = run_variant(ts, torch, case, A, B, C, level, impl, args.steps, args.warmup,
                                                stream, flags)
        if choice:
            level = choice["level"]
        variants[name] = {
            "level": level, "impl": impl, "form": form or "default", "ms_per_step": round(ms, 4),
            "eff_tflops": round(eff_flops / ms / 1e9, 3), "rel_err_vs_L0": err,
            "tol": TOL[level], "ok": err <= TOL[level] if level else True,
            "leaf_kernel_ms": round(kms, 4), "leaf_kernel_tflops": round(kflops / kms / 1e9, 3) if kms else None,
            "leaf_flops_per_launch": kflops, "leaf_launches_per_step": launches / args.steps,
        }
        if choice:
            variants[name]["autotuned"] = choice
        log(f"{name:10s} {ms:8.3f} ms  {eff_flops / ms / 1e9:7.2f} eff TFLOP/s  err {err:.2e}  "
            f"leaf {kms:.3f} ms/launch {kflops / kms / 1e9 if kms else 0:.2f} TFLOP/s")

    cands = [v for v in variants.values() if v["level"] > 0 and v["ok"]]
    head = max(cands, key=lambda v: v["eff_tflops"]) if cands else variants.get("L0_abc")
    classical = variants.get("L0_abc")
    head_name = next(k for k, v in variants.items() if v is head)

    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{workload}_{head_name}.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    achieved = head["leaf_kernel_tflops"]
    lens = "{" + ",".join(str(x) for x in WORKLOADS[workload][0]) + "}"
    return {
        "metric": "effective TFLOP/s (2mnk/t) Strassen L1/L2 vs classical TC, same GPU, at rel-err",
        "value": head["eff_tflops"],
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (splitmix64 uniform[-1,1), seeded index permutations)",
        "config": {
            "workload": f"{workload}: FP64 TC m=n=k={m}, groups of {len(WORKLOADS[workload][0])} indices {lens} in "
                        f"seeded order, seeded index permutations; reported variant {head_name}",
            "m": m, "n": n, "k": k, "idx": f"{case['idx_C']} += {case['idx_A']} * {case['idx_B']}",
            "reported_variant": head_name,
            "rel_err_vs_L0": head["rel_err_vs_L0"],
            "classical_eff_tflops": classical["eff_tflops"] if classical else None,
            "best_abc_eff_tflops": max((v["eff_tflops"] for v in variants.values()
                                        if v["level"] > 0 and v["impl"] == "abc" and v["ok"]), default=None),
            "speedup_vs_classical": round(head["eff_tflops"] / classical["eff_tflops"], 4) if classical else None,
        },
        "roofline": {
            "bound": "mfma",
            "achieved": achieved,
            "peak": PEAK_FP64_MFMA,
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP64_MFMA, 4) if achieved else None,
            "traffic": traffic,
        },
        "variants": variants,
        "cpu_baseline": None,
        "_case": case,
    }


def summary(line):
    """A configuration's JSON line without the fields the headline line already carries."""
    drop = {"n_gpus", "steps", "warmup", "higher_is_better", "scaling", "vs_baseline", "data", "cpu_baseline",
            "_case"}
    return {k: v for k, v in line.items() if k not in drop}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c2", choices=sorted(WORKLOADS) + ["c2p", "c3", "c4", "c5"])
    ap.add_argument("--variants", default="0:abc,1:abc:staged,1:abc:fused,1:abc:staged+atomic,2:abc:staged_all,"
                                          "2:abc:staged_all+atomic,2:abc:staged,2:abc:fused,1:ab,2:ab,1:naive,"
                                          "2:naive,auto",
                    help="level:impl[:form], form = fused (every operand sum in the leaf kernel), staged (the "
                         "outer level's sums staged in HBM) or staged_all (every level's sums staged), optionally "
                         "+atomic (all products in one launch with atomic C updates) or +persist; default: the "
                         "library's choice; `auto`: the autotuner's level, impl and register tile")
    ap.add_argument("--modes", default="", help="c2p / c3: precision:level list (default: all modes)")
    ap.add_argument("--precision-modes", default="fp64:0,ozaki_bf16:1,ozaki_bf16:2,ozaki_int8:1,ozaki_int8:2,"
                                                 "bf16x3:1",
                    help="c2: the precision modes of the headline line's `precision` block")
    ap.add_argument("--no-extras", action="store_true",
                    help="c2: leave out the `precision` block and the c1 / c3 / c4 / c5 summaries")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run one device copy of MIB MiB (a kernel of known traffic for counter scaling)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)

    import torch

    import tblis_strassen as ts

    ts.load()
    torch.cuda.set_device(0)
    stream = torch.cuda.current_stream()
    if args.copy_calibration > 0:
        nbytes = args.copy_calibration << 20
        src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        ts.copy_bytes(dst, src, nbytes, stream=stream)
        torch.cuda.synchronize()
        del src, dst
    if args.workload in ("c2p", "c3", "c4", "c5"):
        if args.workload == "c2p":
            out = bench_c3(ts, torch, args, stream, "c2p")
        else:
            out = {"c3": bench_c3, "c4": bench_c4, "c5": bench_c5}[args.workload](ts, torch, args, stream)
        print(json.dumps(out))
        return
    out = bench_tc(ts, torch, args, stream, args.workload, args.variants)
    case = out.pop("_case")
    torch.cuda.empty_cache()
    if args.workload == "c2" and not args.no_extras:
        # the split-precision target on the same contraction: FP64 classical against the Ozaki modes
        pargs = argparse.Namespace(**vars(args))
        pargs.modes = args.precision_modes
        prec = bench_c3(ts, torch, pargs, stream, "c2p")
        modes = prec["modes"]
        fp64 = modes["fp64_L0"]["eff_tflops"]
        split = {k: v for k, v in modes.items() if v["precision"] == "ozaki_bf16"}
        met = {k: bool(v["eff_tflops"] > fp64 and v["rel_err_vs_fp64"] <= 1e-10) for k, v in split.items()}
        out["precision"] = {
            "workload": "the same contraction (m=n=k=8192), A, B uniform[-1,1); errors vs the FP64 classical "
                        "result; leaf_frac = leaf MFMA rate / that dtype's dense peak",
            "fp64_classical_eff_tflops": fp64,
            "split_bf16_above_fp64_classical_at_1e-10": met,
            "modes": modes,
        }
        torch.cuda.empty_cache()
        out["configs"] = {}
        c1 = bench_tc(ts, torch, args, stream, "c1", "0:abc,1:abc,1:ab,1:naive,2:abc,2:naive")
        c1.pop("_case")
        out["configs"]["c1"] = summary(c1)
        torch.cuda.empty_cache()
        for name, fn in (("c3", bench_c3), ("c4", bench_c4), ("c5", bench_c5)):
            cargs = argparse.Namespace(**vars(args))
            cargs.modes = ""
            out["configs"][name] = summary(fn(ts, torch, cargs, stream))
            torch.cuda.empty_cache()
    if not args.no_cpu_baseline:
        threads = max(1, min(16, host_cpu()["physical_cores"] or 16, len(os.sched_getaffinity(0))))
        out["cpu_baseline"] = cpu_baseline(case, threads)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
