"""Python binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() (as the checker) and by
bench.py's cpu_baseline leg (as the timed CPU port).  Nothing in tblis-strassen_amd/ imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_lib = None

IMPL = {"abc": 0, "ab": 4, "naive": 3}


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    lib = ctypes.CDLL(LIB)
    L = ctypes.POINTER(ctypes.c_long)
    D = ctypes.POINTER(ctypes.c_double)
    s = ctypes.c_char_p
    tens = [ctypes.c_int, L, L, s, D]
    lib.oracle_contract_ref.argtypes = tens * 3 + [ctypes.c_double, ctypes.c_double]
    lib.oracle_tensor_mult.argtypes = [ctypes.c_int] * 4 + tens * 3 + [ctypes.c_double, ctypes.c_double]
    lib.oracle_tensor_mult_blocked.argtypes = [ctypes.c_int] * 2 + tens * 3 + [ctypes.c_double, ctypes.c_double]
    lib.oracle_products.argtypes = [ctypes.c_int, D]
    lib.oracle_describe.argtypes = [ctypes.c_int, L, L, s] * 3 + [L, L, ctypes.c_long]
    lib.oracle_fill_uniform.argtypes = [D, ctypes.c_long, ctypes.c_ulonglong, ctypes.c_int]
    lib.oracle_max_threads.restype = ctypes.c_int
    _lib = lib
    return lib


def _tens(x, idx):
    assert x.dtype == np.float64
    item = x.itemsize
    lens = (ctypes.c_long * max(1, x.ndim))(*x.shape)
    strides = (ctypes.c_long * max(1, x.ndim))(*[st // item for st in x.strides])
    return [x.ndim, lens, strides, idx.encode(), x.ctypes.data_as(ctypes.POINTER(ctypes.c_double))]


def contract_ref(alpha, A, idx_A, B, idx_B, beta, C, idx_C):
    """Brute-force contraction, in place on C (numpy float64 arrays, any strides)."""
    lib = load()
    rc = lib.oracle_contract_ref(*_tens(A, idx_A), *_tens(B, idx_B), *_tens(C, idx_C), alpha, beta)
    if rc:
        raise ValueError("oracle_contract_ref: bad labels/lengths")


def tensor_mult(alpha, A, idx_A, B, idx_B, beta, C, idx_C, level=1, impl="abc", ref_fringe=False, threads=0):
    """Strassen restatement (level 0/1/2, impl abc/ab/naive), in place on C."""
    lib = load()
    impl_c = IMPL[impl] if isinstance(impl, str) else int(impl)
    rc = lib.oracle_tensor_mult(level, impl_c, int(ref_fringe), threads, *_tens(A, idx_A), *_tens(B, idx_B),
                                *_tens(C, idx_C), alpha, beta)
    if rc:
        raise ValueError("oracle_tensor_mult: bad labels/lengths")


def tensor_mult_blocked(alpha, A, idx_A, B, idx_B, beta, C, idx_C, level=1, threads=0):
    """ABC Strassen through the reference's blocked loop nest (packed operand-sum panels, register
    micro-kernel, multi-C update), in place on C; pure contractions only.  The timed CPU baseline."""
    lib = load()
    rc = lib.oracle_tensor_mult_blocked(level, threads, *_tens(A, idx_A), *_tens(B, idx_B), *_tens(C, idx_C),
                                        alpha, beta)
    if rc:
        raise ValueError("oracle_tensor_mult_blocked: bad labels/lengths or not a pure contraction")


def products(level):
    """The oracle's product table: list of (A terms, B terms, C terms), terms = [(subid, coeff)]."""
    lib = load()
    buf = (ctypes.c_double * (49 * 3 * 9))()
    n = lib.oracle_products(level, buf)
    out = []
    for p in range(n):
        prod = []
        for o in range(3):
            q = buf[(p * 3 + o) * 9:(p * 3 + o + 1) * 9]
            prod.append([(int(q[1 + 2 * i]), q[2 + 2 * i]) for i in range(int(q[0]))])
        out.append(tuple(prod))
    return out


def describe(A_shape, A_strides, idx_A, B_shape, B_strides, idx_B, C_shape, C_strides, idx_C, cap=4096):
    """(m, n, k, nbatch, pure) and the scatter vectors a_m, a_k, b_k, b_n, c_m, c_n (un-transposed)."""
    lib = load()
    args = []
    for sh, st, ix in ((A_shape, A_strides, idx_A), (B_shape, B_strides, idx_B), (C_shape, C_strides, idx_C)):
        args += [len(sh), (ctypes.c_long * max(1, len(sh)))(*sh), (ctypes.c_long * max(1, len(sh)))(*st),
                 ix.encode()]
    dims = (ctypes.c_long * 5)()
    scat = (ctypes.c_long * (6 * cap))()
    if lib.oracle_describe(*args, dims, scat, cap):
        raise ValueError("oracle_describe failed")
    d = list(dims)
    sizes = [d[0], d[2], d[2], d[1], d[0], d[1]]
    vecs = [np.array(scat[i * cap:i * cap + min(cap, sizes[i])], dtype=np.int64) for i in range(6)]
    return d, vecs


def fill_uniform(n, seed, emax=0):
    lib = load()
    out = np.empty(n, dtype=np.float64)
    lib.oracle_fill_uniform(out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n, seed, emax)
    return out


def max_threads():
    return load().oracle_max_threads()
