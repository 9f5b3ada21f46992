#!/usr/bin/env python
"""Shape-file driver: run a list of contractions through the library's C-ABI and report the
reference harness's columns (test/test_strassen.cxx:45-66, 299-413, 493-530).

Usage: python tblis-strassen_amd/shape_driver.py -file SHAPES [-niter R] [-level L] [-impl I] [-seed S]
                                                 [-prec double|float] [-notrash]
A shape line is `idxC idxA idxB & a:len;b:len;...;` ('#' lines are skipped, the first empty line ends
the list).  -impl: 1 = ABC, 2 = AB, 3 = Naive (as the reference's CLI).  A and B are uniform [0,1)
(the reference fills with drand48), C and C_ref start at zero and every run adds A.B (beta = 1).
Columns: m n k, classical seconds, Strassen seconds (minimum over R runs), GFLOP, GFLOP/s of each, and
the reference's error ||C - C_ref||^2 / ||C_ref||^2 (squared, as compare_tensor reports it).
Times are device times of one synchronous call each, with the device caches flushed between runs
(a 512 MiB write, the counterpart of trashCache) unless -notrash is given.
"""
import argparse
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

IMPL_CLI = {1: "abc", 2: "ab", 3: "naive"}


class ShapeError(ValueError):
    pass


def parse_lengths(text):
    """`a:480;b:480;` -> {'a': 480, 'b': 480}; every entry ends with ';' and has its ':' at position 1."""
    lengths = {}
    s = text
    while ";" in s:
        token, s = s.split(";", 1)
        if token.find(":") != 1:
            raise ShapeError(f"bad length entry {token!r}")
        lengths[token[0]] = int(token[2:])
    return lengths


def parse_shape_lines(lines):
    """Yield (idx_C, idx_A, idx_B, lengths) for each shape line."""
    for line in lines:
        line = line.rstrip("\n")
        if not line:
            break
        if line.startswith("#"):
            continue
        parts = line.split()
        if len(parts) < 5 or parts[3] != "&":
            raise ShapeError(f"bad shape line {line!r}")
        idx_C, idx_A, idx_B = parts[0], parts[1], parts[2]
        lengths = parse_lengths(parts[4])
        for c in idx_A + idx_B + idx_C:
            if c not in lengths:
                raise ShapeError(f"label {c!r} has no length in {line!r}")
        yield idx_C, idx_A, idx_B, lengths


def group_sizes(idx_A, idx_B, idx_C, lengths):
    """(m, n, k, total label-space size): the extents of the AC, BC and AB label groups."""
    def prod(labels):
        p = 1
        for c in labels:
            p *= lengths[c]
        return p
    m = prod([c for c in idx_A if c in idx_C and c not in idx_B])
    n = prod([c for c in idx_B if c in idx_C and c not in idx_A])
    k = prod([c for c in idx_A if c in idx_B and c not in idx_C])
    return m, n, k, prod(sorted(lengths))


def run_line(ts, torch, idx_C, idx_A, idx_B, lengths, niter, level, impl, dtype, seed, trash, keep=False):
    """Time one shape line (classical, then the chosen level and impl); keep=True also returns the flat
    device buffers of A, B, C (Strassen) and C_ref (classical) under "buffers"."""
    def colmajor(shape):
        st, s = [], 1
        for L in shape:
            st.append(s)
            s *= L
        return st

    def dev(labels, fill_seed=None):
        shape = [lengths[c] for c in labels]
        n = 1
        for L in shape:
            n *= L
        buf = torch.zeros(max(n, 1), dtype=dtype, device="cuda")
        if fill_seed is not None:
            ts.fill_uniform(buf, fill_seed)
            buf.add_(1.0).mul_(0.5)  # uniform [0,1), the reference's drand48 range
        return buf.as_strided(shape, colmajor(shape)), buf

    A, abuf = dev(idx_A, seed * 2 + 1)
    B, bbuf = dev(idx_B, seed * 2 + 2)
    C, cbuf = dev(idx_C)
    Cref, crbuf = dev(idx_C)

    def timed(fn):
        best = float("inf")
        for _ in range(niter):
            if trash is not None:
                trash.fill_(0.0)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
        return best

    dt = timed(lambda: ts.tensor_mult(1.0, A, idx_A, B, idx_B, 1.0, Cref, idx_C, level=0, impl="abc"))
    dt2 = timed(lambda: ts.tensor_mult(1.0, A, idx_A, B, idx_B, 1.0, C, idx_C, level=level, impl=impl))
    diff = (cbuf.double() - crbuf.double()).pow(2).sum().item()
    ref = crbuf.double().pow(2).sum().item()
    err2 = diff / (ref if ref != 0.0 else 1.0)
    m, n, k, ntot = group_sizes(idx_A, idx_B, idx_C, lengths)
    gflops = 2.0 * ntot * 1e-9
    out = dict(m=m, n=n, k=k, dt=dt, dt2=dt2, gflops=gflops, rate=gflops / dt, rate2=gflops / dt2, err2=err2)
    if keep:
        out["buffers"] = dict(A=abuf, B=bbuf, C=cbuf, Cref=crbuf)
    return out


def format_row(r):
    return (f"{r['m']:5d} {r['n']:5d} {r['k']:5d}\t {r['dt']:10.3e} {r['dt2']:10.3e} {r['gflops']:8.3e}\t "
            f"{r['rate']:7.2f} {r['rate2']:7.2f}\t {r['err2']:5.2f}")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter,
                                 prefix_chars="-")
    ap.add_argument("-file", required=True)
    ap.add_argument("-niter", type=int, default=3)
    ap.add_argument("-level", type=int, default=1, choices=(0, 1, 2))
    ap.add_argument("-impl", type=int, default=1, choices=(1, 2, 3))
    ap.add_argument("-seed", type=int, default=0)
    ap.add_argument("-prec", default="double", choices=("double", "float"))
    ap.add_argument("-notrash", action="store_true")
    args = ap.parse_args(argv)
    if args.niter < 0:
        args.niter = 3

    import torch

    import tblis_strassen as ts

    ts.load()
    dtype = torch.float64 if args.prec == "double" else torch.float32
    trash = None if args.notrash else torch.empty(64 << 20, dtype=torch.float64, device="cuda")
    with open(args.file) as f:
        shapes = list(parse_shape_lines(f))
    for idx_C, idx_A, idx_B, lengths in shapes:
        r = run_line(ts, torch, idx_C, idx_A, idx_B, lengths, args.niter, args.level, IMPL_CLI[args.impl], dtype,
                     args.seed, trash)
        print(format_row(r), flush=True)


if __name__ == "__main__":
    main()
