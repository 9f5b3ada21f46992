"""CPU-only checks of the HIP library's host side: it loads, exports every symbol the headers declare,
and its index analysis / Strassen plan agree with the oracle (no device calls are made here)."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

import cases
import oracle
import tblis_strassen as ts

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"^\s*(?:const\s+)?(?:void|int|char)\s*\*?\s*(\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
        if re.search(r"extern\s+const\s+tblis_comm\s*\*\s*const\s+tblis_single", text):
            names.add("tblis_single")
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = ts.load()
    names = _declared_symbols()
    assert "stra_tblis_tensor_mult" in names and "mistra_tensor_mult" in names and len(names) > 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_combine_form_switch_returns_previous_setting():
    lib = ts.load()
    lib.mistra_debug_set_combine_form.restype = ctypes.c_int
    lib.mistra_debug_set_combine_form.argtypes = [ctypes.c_int]
    old = lib.mistra_debug_set_combine_form(1)
    assert lib.mistra_debug_set_combine_form(7) == 1      # any non-zero value means the per-sub-block form
    assert lib.mistra_debug_set_combine_form(old) == 1


def test_cxx_impl_global_is_exported():
    # tblis::internal::impl keeps its C++ name for drop-in users of the reference's global
    lib = ts.load()
    assert hasattr(lib, "_ZN5tblis8internal4implE")
    old = ts.get_impl()
    ts.set_impl("naive")
    assert ts.get_impl() == 3 and ctypes.c_int.in_dll(lib, "_ZN5tblis8internal4implE").value == 3
    ts.set_impl(old)


def _debug_plan(case, level, flags=0, strides=None, cap=8192):
    lib = ts.load()
    T = ctypes.POINTER(ts.TblisTensor)
    lib.mistra_debug_plan.restype = ctypes.c_int
    lib.mistra_debug_plan.argtypes = [T, ctypes.c_char_p, T, ctypes.c_char_p, T, ctypes.c_char_p, ctypes.c_int,
                                      ctypes.c_uint, ctypes.POINTER(ctypes.c_longlong),
                                      ctypes.POINTER(ctypes.c_longlong), ctypes.c_longlong,
                                      ctypes.POINTER(ctypes.c_double), ctypes.c_int]
    tensors = []
    for x in "ABC":
        shape = case[f"shape_{x}"]
        st = strides[x] if strides else cases.colmajor_strides(shape)
        t = ts.TblisTensor()
        t.type = ts.TYPE_DOUBLE
        t.ndim = len(shape)
        t._len = (ctypes.c_ssize_t * max(1, len(shape)))(*shape)
        t._stride = (ctypes.c_ssize_t * max(1, len(shape)))(*st)
        t.len = ctypes.cast(t._len, ctypes.POINTER(ctypes.c_ssize_t))
        t.stride = ctypes.cast(t._stride, ctypes.POINTER(ctypes.c_ssize_t))
        t.data = None
        tensors.append(t)
    dims = (ctypes.c_longlong * 16)()
    scat = (ctypes.c_longlong * (6 * cap))()
    prods = (ctypes.c_double * (64 * 40))()
    np_ = lib.mistra_debug_plan(ctypes.byref(tensors[0]), case["idx_A"].encode(), ctypes.byref(tensors[1]),
                                case["idx_B"].encode(), ctypes.byref(tensors[2]), case["idx_C"].encode(), level, flags,
                                dims, scat, cap, prods, 64)
    return np_, list(dims), [np.array(scat[i * cap:(i + 1) * cap]) for i in range(6)], np.array(prods[:np_ * 40] if np_ > 0 else [])


PLAN_CASES = [
    dict(m=(6,), n=(5,), k=(7,)),
    dict(m=(4, 4), n=(4, 4), k=(4, 4)),
    dict(m=(2, 2, 4), n=(2, 2, 4), k=(2, 2, 4)),
    dict(m=(3, 5), n=(2, 7), k=(9,)),
    dict(m=(4,), n=(8,), k=(4,), batch=(3,)),
]


@pytest.mark.parametrize("ci", range(len(PLAN_CASES)))
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_index_analysis_matches_oracle(ci, seed):
    case = cases.make_case(1000 * seed + ci, **PLAN_CASES[ci])
    nprod, dims, vecs, _ = _debug_plan(case, 0)
    assert nprod == 1
    m, n, k, nb, pure, transposed = dims[:6]
    sA, sB, sC = (cases.colmajor_strides(case[f"shape_{x}"]) for x in "ABC")
    (odims, ovecs) = oracle.describe(case["shape_A"], sA, case["idx_A"], case["shape_B"], sB, case["idx_B"],
                                     case["shape_C"], sC, case["idx_C"])
    am, ak, bk, bn, cm, cn = ovecs
    if transposed:
        want = [bn, bk, ak, am, cn, cm]
        assert (m, n, k) == (odims[1], odims[0], odims[2])
    else:
        want = [am, ak, bk, bn, cm, cn]
        assert (m, n, k) == (odims[0], odims[1], odims[2])
    assert nb == odims[3] and pure == 1
    # the scatter vectors hold the oracle's offset pairs per group (a group may take another,
    # documented index order: the one that gives its operand the smallest leading stride)
    for i, j in ((0, 4), (1, 2), (3, 5)):
        nn = len(want[i])
        got_pairs = sorted(zip(vecs[i][:nn].tolist(), vecs[j][:nn].tolist()))
        assert got_pairs == sorted(zip(want[i].tolist(), want[j].tolist()))


@pytest.mark.parametrize("level", [1, 2])
def test_product_table_matches_oracle(level):
    case = cases.make_case(7, m=(16,), n=(16,), k=(16,))
    nprod, dims, _, prods = _debug_plan(case, level)
    assert nprod == 7 ** level and dims[10] == 0  # no fringe passes
    d = 2 ** level
    ms, ns, ks = dims[6:9]
    table = oracle.products(level)
    touched = set()
    for q, (A, B, C) in enumerate(table):
        o = prods[q * 40:(q + 1) * 40]
        assert (o[0], o[1], o[2]) == (len(A), len(B), len(C))
        for si, (terms, ext_r, ext_c) in enumerate(((A, ms, ks), (B, ks, ns), (C, ms, ns))):
            for t, (sub, coef) in enumerate(terms):
                r = o[4 + si * 12 + t * 3:4 + si * 12 + t * 3 + 3]
                assert (r[0], r[1], r[2]) == ((sub // d) * ext_r, (sub % d) * ext_c, coef)
        mask = 0
        for t, (sub, _) in enumerate(C):
            if sub not in touched:
                mask |= 1 << t
                touched.add(sub)
        assert int(o[3]) == mask


@pytest.mark.parametrize("level", [1, 2])
def test_fringe_passes_cover_remainder(level):
    case = cases.make_case(9, m=(13,), n=(10,), k=(7,))
    nprod, dims, _, prods = _debug_plan(case, level)
    d = 2 ** level
    m, n, k = dims[:3]
    ms, ns, ks, ncore, nfringe = dims[6:11]
    assert (ms, ns, ks) == (m // d, n // d, k // d)
    assert nfringe == sum([m % d != 0, n % d != 0, k % d != 0])
    _, dims_rf, _, _ = _debug_plan(case, level, flags=ts.FLAG_REF_FRINGE)
    assert dims_rf[10] == 0


def test_transpose_when_c_is_row_major():
    # C stored with its n index unit-stride -> the plan computes C^T = B^T A^T
    case = cases.make_case(3, m=(8,), n=(8,), k=(8,), permute=False)
    strides = {"A": [1, 8], "B": [1, 8], "C": [8, 1]}
    _, dims, vecs, _ = _debug_plan(case, 1, strides=strides)
    assert dims[5] == 1
    assert np.array_equal(vecs[4][:8], np.arange(8))  # rows of C^T are unit-stride


def test_run_strides_bruteforce():
    lib = ts.load()
    rng = np.random.default_rng(0)
    for trial in range(50):
        n = int(rng.integers(1, 60))
        run = int(rng.choice([4, 16, 32]))
        # piecewise-affine vectors, like the scatter of a multi-index group
        L0 = int(rng.integers(1, 9))
        s0, s1 = int(rng.integers(1, 5)), int(rng.integers(5, 50))
        v = np.array([(p % L0) * s0 + (p // L0) * s1 for p in range(n)], dtype=np.int64)
        if trial % 5 == 0:
            v = rng.permutation(v)
        out = (ctypes.c_longlong * n)()
        lib.mistra_debug_run_strides.argtypes = [ctypes.POINTER(ctypes.c_longlong), ctypes.c_longlong, ctypes.c_int,
                                                 ctypes.POINTER(ctypes.c_longlong)]
        lib.mistra_debug_run_strides(v.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), n, run, out)
        for p in range(n):
            cnt = min(run, n - p)
            if cnt <= 1:
                continue
            d = v[p + 1] - v[p]
            affine = d != 0 and all(v[p + j] == v[p] + j * d for j in range(cnt))
            assert out[p] == (d if affine else 0), (trial, p)


def test_label_length_mismatch_is_an_error():
    case = cases.make_case(1, m=(4,), n=(4,), k=(4,), permute=False)
    case = dict(case)
    case["shape_B"] = [5, 4]  # k has length 4 in A, 5 in B
    nprod, _, _, _ = _debug_plan(case, 1)
    assert nprod == -1  # MISTRA_ERR_ARG


def test_binding_writes_scalars_and_strides():
    # tblis_tensor.scalar holds alpha/beta as the tensor's own type (src/util/basic_types.h:341-419)
    x = np.zeros((3, 4), dtype=np.float64, order="F")
    t = ts.make_tensor(x, 2.5)
    raw = bytes(t)
    assert np.frombuffer(raw[8:16], dtype=np.float64)[0] == 2.5
    assert t.type == ts.TYPE_DOUBLE and t.ndim == 2 and [t.len[0], t.len[1]] == [3, 4]
    assert [t.stride[0], t.stride[1]] == [1, 3]
    t32 = ts.make_tensor(np.zeros((2, 2), dtype=np.float32), -0.5)
    assert np.frombuffer(bytes(t32)[8:12], dtype=np.float32)[0] == -0.5


def test_struct_layouts_match_the_header(tmp_path):
    # the ctypes images used by the binding and the C structs of include/tblis_strassen.h agree byte for byte
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "tblis_strassen.h"\n'
        "int main(void) {\n"
        '  printf("%zu %zu %zu %zu %zu %zu\\n", sizeof(tblis_tensor), offsetof(tblis_tensor, scalar),\n'
        "         offsetof(tblis_tensor, data), offsetof(tblis_tensor, ndim), offsetof(tblis_tensor, len),\n"
        "         offsetof(tblis_tensor, stride));\n"
        '  printf("%zu %zu %zu %zu %zu\\n", sizeof(tblis_matrix), offsetof(tblis_matrix, scalar),\n'
        "         offsetof(tblis_matrix, data), offsetof(tblis_matrix, m), offsetof(tblis_matrix, cs));\n"
        "  return 0; }\n")
    exe = tmp_path / "layout"
    import subprocess
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    T, M = ts.TblisTensor, ts.TblisMatrix
    assert [int(x) for x in out[0].split()] == [ctypes.sizeof(T), T.scalar.offset, T.data.offset, T.ndim.offset,
                                                T.len.offset, T.stride.offset]
    assert [int(x) for x in out[1].split()] == [ctypes.sizeof(M), M.scalar.offset, M.data.offset, M.m.offset,
                                                M.cs.offset]
    assert ctypes.sizeof(T) == 56 and ctypes.sizeof(M) == 64


def test_autotuner_analytic_model():
    # host-only choice (no device data touched): big cubic FP64 contractions get 128x128 tiles and a
    # Strassen level only where the model predicts a gain; tiny ones stay classical
    import bench

    case = bench.make_workload("c2")
    A, B, C = (cases.host_tensor(case[f"shape_{x}"], np.zeros(1)) if False else
               np.lib.stride_tricks.as_strided(np.zeros(1), shape=case[f"shape_{x}"],
                                               strides=[0] * len(case[f"shape_{x}"])) for x in "ABC")
    level, impl, tile = ts.autotune(A, case["idx_A"], B, case["idx_B"], C, case["idx_C"])
    assert level in (0, 1, 2) and impl in (0, 3, 4) and tile == 0  # TILE_128
    small = cases.make_case(3, m=(8,), n=(8,), k=(8,))
    Z = [np.zeros(int(np.prod(small[f"shape_{x}"]))).reshape(small[f"shape_{x}"], order="F") for x in "ABC"]
    level, impl, tile = ts.autotune(Z[0], small["idx_A"], Z[1], small["idx_B"], Z[2], small["idx_C"])
    assert level == 0 and tile == 1  # TILE_64


def test_autotuner_model_picks_measured_fastest_at_c2():
    # with c2's real (column-major, seeded-permutation) layout the model's costs follow the measured
    # ones (classical 15.3 ms, L1 ABC staged 14.5, L2 ABC staged 14.6, L2 Naive 13.5-14.2 on MI355X) and
    # it picks L2 Naive with 64 x 128 tiles, on average the fastest of that launch's tiles (they trade
    # places by up to 4% from one process to the next); only the descriptors are read
    import bench

    case = bench.make_workload("c2")
    buf = np.zeros(1)
    A, B, C = (np.lib.stride_tricks.as_strided(buf, shape=case[f"shape_{x}"],
                                               strides=[8 * s for s in cases.colmajor_strides(case[f"shape_{x}"])])
               for x in "ABC")
    choice = ts.autotune(A, case["idx_A"], B, case["idx_B"], C, case["idx_C"])
    assert choice == (2, ts.IMPL["naive"], 3), choice


def test_autotuner_model_picks_measured_fastest_at_c3():
    # m = n = 16384, k = 1024 (2 x 2 x 2 indices, column-major): classical with 64 x 128 tiles measured
    # 8.11 / 8.23 ms against 8.28 / 8.32 with 128 x 128 on MI355X (two runs; 128 x 64 8.93, every Strassen
    # form slower): at a 1024-deep leaf the wide tile's three resident workgroups per CU hide more of each
    # tile's start and end
    case = cases.make_case(4242, m=(128, 128), n=(128, 128), k=(32, 32))
    buf = np.zeros(1)
    A, B, C = (np.lib.stride_tricks.as_strided(buf, shape=case[f"shape_{x}"],
                                               strides=[8 * s for s in cases.colmajor_strides(case[f"shape_{x}"])])
               for x in "ABC")
    choice = ts.autotune(A, case["idx_A"], B, case["idx_B"], C, case["idx_C"])
    assert choice == (0, 0, 3), choice


def test_abc_stage_counts():
    # first-fit-decreasing packing by C targets: 4 launches at one level, 16 at two (optimal for 49
    # products of equal size: every stage but one holds an even number of products)
    case = cases.make_case(8, m=(32,), n=(32,), k=(32,))
    for level, want in ((0, 1), (1, 4), (2, 16)):
        _, dims, _, _ = _debug_plan(case, level, flags=ts.FLAG_FUSED_SUMS)
        assert dims[14] == want


def test_outer_staged_choice():
    # the cost model stages operand sums of 1- and 2-level ABC plans at 8192^3 (sum traffic ~ n^2
    # against the slower multi-term leaf launches ~ n^3); the flags force a form (staged depth 0, 1 or
    # the level); classical and batched plans never stage
    big = cases.make_case(9, m=(8192,), n=(8192,), k=(8192,))
    small = cases.make_case(9, m=(64,), n=(64,), k=(64,))
    batched = cases.make_case(9, m=(64,), n=(64,), k=(64,), batch=(2,))
    assert _debug_plan(big, 2)[1][13] >= 1
    assert _debug_plan(big, 2, flags=ts.FLAG_FUSED_SUMS)[1][13] == 0
    assert _debug_plan(big, 1)[1][13] == 1
    assert _debug_plan(big, 0)[1][13] == 0
    assert _debug_plan(small, 2, flags=ts.FLAG_FUSED_SUMS)[1][13] == 0
    assert _debug_plan(small, 2, flags=ts.FLAG_OUTER_STAGED)[1][13] == 1
    assert _debug_plan(small, 2, flags=ts.FLAG_ALL_STAGED)[1][13] == 2
    assert _debug_plan(small, 1, flags=ts.FLAG_ALL_STAGED)[1][13] == 1
    assert _debug_plan(batched, 2, flags=ts.FLAG_OUTER_STAGED)[1][13] == 0
    assert _debug_plan(batched, 2, flags=ts.FLAG_OUTER_STAGED)[1][15] == 0


def test_packed_classical_choice():
    # level 0: A and B are packed into K-fast copies when an operand is not K-fast and the model says
    # so (8192^3 with a column-major B, i.e. N-fast after the plan's orientation), never when both are
    # K-fast already, and on request for small shapes
    big = cases.make_case(9, m=(8192,), n=(8192,), k=(8192,))
    strides = {"A": [8192, 1], "B": [8192, 1], "C": [1, 8192]}  # A: k fast; B[k][n]: n fast; C column-major
    assert _debug_plan(big, 0, strides=strides)[1][15] == 1
    kfast = {"A": [8192, 1], "B": [1, 8192], "C": [1, 8192]}
    assert _debug_plan(big, 0, strides=kfast)[1][15] == 0
    small = cases.make_case(9, m=(64,), n=(64,), k=(64,))
    assert _debug_plan(small, 0, flags=ts.FLAG_OUTER_STAGED)[1][15] == 1
    assert _debug_plan(small, 0, flags=ts.FLAG_FUSED_SUMS)[1][15] == 0


def test_tune_candidates_cover_four_tiles_where_launches_allow():
    # the rectangular tiles serve single-term launches only: classical, Naive's dense GEMM, staged ABC;
    # AB (operand sums formed in the leaf kernel) is offered with the square tiles alone
    import bench

    case = bench.make_workload("c2")
    buf = np.zeros(1)
    A, B, C = (np.lib.stride_tricks.as_strided(buf, shape=case[f"shape_{x}"],
                                               strides=[8 * s for s in cases.colmajor_strides(case[f"shape_{x}"])])
               for x in "ABC")
    cands = ts.autotune_candidates(A, case["idx_A"], B, case["idx_B"], C, case["idx_C"])
    got = {(lv, im, ti) for lv, im, ti, _ in cands}
    naive, ab = ts.IMPL["naive"], ts.IMPL["ab"]
    assert {(0, 0, t) for t in range(4)} <= got
    assert {(2, naive, t) for t in range(4)} <= got
    assert not any(lv > 0 and im == ab and ti >= 2 for lv, im, ti in got)
    ms = [m for *_, m in cands]
    assert ms == sorted(ms) and all(m > 0 for m in ms)
