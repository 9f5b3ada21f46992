#!/usr/bin/env python
"""Benchmark: effective TFLOP/s (2mnk/t) of Strassen tensor contraction (L=1, L=2) against the classical
(L=0) contraction on the same MI355X, FP64, at the stated relative error.

Workload (one GPU): m = n = k = 8192; each of the m, n, k groups has 3 indices of lengths {16,16,32} in
seeded order; each tensor's index order is a seeded permutation; A, B ~ uniform[-1,1) (splitmix64,
generated in HBM); C starts at zero and every step is C += A.B (beta = 1, as the reference's
test_strassen does).  A step is one contraction call; inputs are resident in HBM before timing.

Variants: L0 (classical MFMA TC), L1/L2 x {ABC, AB, Naive}.  Errors are ||C - C_L0||_F / ||C_L0||_F of a
single call from C = 0.  The reported `value` is the fastest Strassen variant (L1 or L2, any impl) that
meets its tolerance (1e-12 at L1, 5e-12 at L2); every variant, ABC included, is listed under `variants`.

roofline: FP64 MFMA bound of the leaf kernel of the reported variant: achieved = leaf flops per launch /
average launch duration (HIP events recorded by the library on the launch stream); traffic comes from the
committed counter summary profiles/traffic_<workload>.json when present.
cpu_baseline: the CPU oracle (a port of the reference algorithm) on a bounded sample of the workload.
"""
import argparse
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


def run_variant(ts, torch, case, A, B, C, level, impl, steps, warmup, stream):
    """Time `steps` calls after `warmup` calls; returns (ms/step, kernel ms/launch, launches, leaf flops/launch)."""
    for _ in range(warmup):
        ts.tensor_mult(1.0, A, case["idx_A"], B, case["idx_B"], 1.0, C, case["idx_C"], level=level, impl=impl,
                       stream=stream)
    torch.cuda.synchronize()
    ts.kernel_timing(True)
    ts.kernel_timing_read(reset=True)
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record(stream)
    for _ in range(steps):
        ts.tensor_mult(1.0, A, case["idx_A"], B, case["idx_B"], 1.0, C, case["idx_C"], level=level, impl=impl,
                       stream=stream)
    stop.record(stream)
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop) / steps
    kms, launches, kflops = ts.kernel_timing_read(reset=True)
    ts.kernel_timing(False)
    return ms, kms / max(1, launches), launches, kflops / max(1, launches)


def cpu_baseline(threads):
    """The oracle's 1-level ABC Strassen on a bounded sample: the same index structure at 4096^3."""
    import numpy as np

    import cases
    import oracle

    oracle.load()
    lens = (16, 16, 16)
    case = cases.make_case(777, m=lens, n=lens, k=lens)
    m, n, k = cases.mnk(case)
    a, b = oracle.fill_uniform(m * k, 1), oracle.fill_uniform(k * n, 2)
    c = np.zeros(m * n)
    A, B, C = (cases.host_tensor(case[f"shape_{x}"], v) for x, v in zip("ABC", (a, b, c)))
    t0 = time.perf_counter()
    oracle.tensor_mult(1.0, A, case["idx_A"], B, case["idx_B"], 1.0, C, case["idx_C"], level=1, impl="abc",
                       threads=threads)
    dt = time.perf_counter() - t0
    return {
        "value": round(2.0 * m * n * k / dt / 1e12, 5),
        "unit": "TFLOP/s",
        "cores": threads,
        "kind": "port",
        "sample": f"oracle 1-level ABC Strassen, FP64, m=n=k=4096 (3 indices of 16 per group, seeded "
                  f"permutations), one call in {dt:.2f} s with {threads} OpenMP threads",
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c2", choices=sorted(WORKLOADS))
    ap.add_argument("--variants", default="0:abc,1:abc,2:abc,1:ab,2:ab,1:naive,2:naive")
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
    case = make_workload(args.workload)
    import cases

    m, n, k = cases.mnk(case)
    eff_flops = 2.0 * m * n * k
    log(f"workload {args.workload}: C[{case['idx_C']}] += A[{case['idx_A']}] B[{case['idx_B']}]  m=n=k={m}")

    A, abuf = device_tensor(case["shape_A"], torch.float64, torch)
    B, bbuf = device_tensor(case["shape_B"], torch.float64, torch)
    ts.fill_uniform(abuf, 1, stream=stream)
    ts.fill_uniform(bbuf, 2, stream=stream)
    Cref, crefbuf = device_tensor(case["shape_C"], torch.float64, torch)
    crefbuf.zero_()
    ts.tensor_mult(1.0, A, case["idx_A"], B, case["idx_B"], 1.0, Cref, case["idx_C"], level=0, stream=stream)
    C, cbuf = device_tensor(case["shape_C"], torch.float64, torch)

    variants = {}
    for spec in args.variants.split(","):
        lv, impl = spec.split(":")
        level = int(lv)
        name = f"L{level}_{impl}"
        cbuf.zero_()
        ts.tensor_mult(1.0, A, case["idx_A"], B, case["idx_B"], 1.0, C, case["idx_C"], level=level, impl=impl,
                       stream=stream)
        err = ts.rel_error(cbuf, crefbuf, stream=stream)
        cbuf.zero_()
        ms, kms, launches, kflops = run_variant(ts, torch, case, A, B, C, level, impl, args.steps, args.warmup,
                                                stream)
        variants[name] = {
            "level": level, "impl": impl, "ms_per_step": round(ms, 4),
            "eff_tflops": round(eff_flops / ms / 1e9, 3), "rel_err_vs_L0": err,
            "tol": TOL[level], "ok": err <= TOL[level] if level else True,
            "leaf_kernel_ms": round(kms, 4), "leaf_kernel_tflops": round(kflops / kms / 1e9, 3) if kms else None,
            "leaf_flops_per_launch": kflops, "leaf_launches_per_step": launches / args.steps,
        }
        log(f"{name:10s} {ms:8.3f} ms  {eff_flops / ms / 1e9:7.2f} eff TFLOP/s  err {err:.2e}  "
            f"leaf {kms:.3f} ms/launch {kflops / kms / 1e9 if kms else 0:.2f} TFLOP/s")

    cands = [v for v in variants.values() if v["level"] > 0 and v["ok"]]
    head = max(cands, key=lambda v: v["eff_tflops"]) if cands else variants.get("L0_abc")
    classical = variants.get("L0_abc")
    head_name = f"L{head['level']}_{head['impl']}"

    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{args.workload}_{head_name}.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    achieved = head["leaf_kernel_tflops"]
    out = {
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
            "workload": f"{args.workload}: FP64 TC m=n=k={m}, groups of 3 indices {{16,16,32}} in seeded order, "
                        f"seeded index permutations; reported variant {head_name}",
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
    }
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(threads)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
