"""The C++ template surface (include/tblis_strassen.hpp) compiled into a caller program
(tests/cpp/stra_mult_caller.cpp): tblis::stra_mult<T>, stra_mult_2level<T> and mult<T> with
tblis::internal::impl set to BLIS_BASED, STRA_AB and STRA_NAIVE in turn (src/stra_iface/3t/stra_mult.h:26-62,
stra_mult_2level.h:26-62, src/iface/3t/mult.h:26-62), the extension and the matrix overloads, each bit-exact
against a brute-force contraction of small integers computed by the program itself.

CPU: the program builds against the header (g++, no HIP device code).  GPU: it runs and every check passes.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")
BIN = os.path.join(CPP, "stra_mult_caller")


def test_caller_builds_against_the_header():
    r = subprocess.run(["make", "-s", "-C", CPP], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert os.path.exists(BIN)


@pytest.mark.gpu
def test_caller_runs_bit_exact():
    assert os.path.exists(BIN), "tests/cpp/stra_mult_caller is built by __graft_entry__.build()"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL OK" in r.stdout
    assert r.stdout.count(" ok ") == 36
