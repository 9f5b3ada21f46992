"""CPU: the persistent batch launch's job queue, built on the host alone and replayed in queue order
(mistra_debug_batch_check): every tile's pack and round waits must be met by the jobs before it, every
pack chunk must fall in its panel, and every product's pack count must come out exact.  Covers the
batches the GPU tests run (seeded uneven shapes at L0/1/2, fringes, FP32/FP64)."""
import ctypes

import numpy as np
import pytest

import cases
import tblis_strassen as ts


def check(batch, dtype=np.float64, level=1):
    lib = ts.load()
    ents = []
    keep = []
    for case in batch:
        arrs = [np.zeros(int(np.prod(case[f"shape_{x}"])), dtype=dtype) for x in "ABC"]
        tens = [cases.host_tensor(case[f"shape_{x}"], a.astype(np.float64)).astype(dtype) if False else
                np.lib.stride_tricks.as_strided(a, shape=tuple(case[f"shape_{x}"]),
                                                strides=tuple(s * a.itemsize for s in cases.colmajor_strides(case[f"shape_{x}"])))
                for x, a in zip("ABC", arrs)]
        keep.append((arrs, tens))
        ents.append(dict(A=tens[0], idx_A=case["idx_A"], B=tens[1], idx_B=case["idx_B"], C=tens[2], idx_C=case["idx_C"],
                         beta=0.5, level=level, impl="abc"))
    b = ts.Batch(ents)
    lib.mistra_debug_batch_check.argtypes = [ctypes.c_int, ctypes.POINTER(ts.Contraction), ctypes.c_char_p, ctypes.c_int]
    msg = ctypes.create_string_buffer(512)
    rc = lib.mistra_debug_batch_check(b.count, b.arr, msg, 512)
    return rc, msg.value.decode()


@pytest.mark.parametrize("level", [0, 1, 2])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_queue_of_seeded_batches(level, dtype):
    for seed, count, lo, hi in ((5, 16, 32, 700), (11, 24, 16, 600), (2024, 64, 64, 4096)):
        rc, msg = check(cases.c4_batch(seed, count=count, lo=lo, hi=hi), dtype, level)
        assert rc == 0, (seed, msg)


def test_queue_of_uneven_shapes():
    shapes = [dict(m=(13, 7), n=(11, 5), k=(9, 6)), dict(m=(64,), n=(48,), k=(40,)), dict(m=(33, 4), n=(70,), k=(3, 25)),
              dict(m=(5,), n=(260,), k=(7, 9))]
    for level in (0, 1, 2):
        rc, msg = check([cases.make_case(70 + i, **g) for i, g in enumerate(shapes)], np.float64, level)
        assert rc == 0, (level, msg)
