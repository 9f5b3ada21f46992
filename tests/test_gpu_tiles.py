"""Every register tile of the leaf kernel (128x128, 64x64 and the rectangular 128x64 / 64x128 that serve
single-term launches) on the same parity bars as the default choice, forced through
mistra_debug_set_tile:
  * small integer inputs bit-exact against numpy.einsum at L0 / L1 / L2 for ABC with staged sums (single-
    term stage launches), Naive (single-term dense GEMM) and AB (multi-term: the plan falls back to a
    square tile), on shapes with fringes, a batch index and leaf extents that are not tile multiples;
  * uniform FP64 inputs within 1e-13 of the oracle (same level and impl), FP32 within 1e-5 of FP64 einsum;
  * the autotuner's candidates include the rectangular tiles, and its measured choice runs.
"""
import numpy as np
import pytest

import cases
import tblis_strassen as ts

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

from test_gpu_parity import dev, einsum, inputs, run_gpu, run_oracle, say  # noqa: E402

TILES = {0: "128x128", 1: "64x64", 2: "128x64", 3: "64x128"}
CASES = [
    dict(m=(64, 4), n=(8, 40), k=(96,)),
    dict(m=(160,), n=(136,), k=(200,)),
    dict(m=(12, 20), n=(36,), k=(6, 12)),
    dict(m=(24,), n=(40,), k=(8,), batch=(3,)),
    dict(m=(300,), n=(72,), k=(64,)),
]
FORMS = [("abc", ts.FLAG_OUTER_STAGED), ("abc", ts.FLAG_ALL_STAGED), ("naive", 0), ("ab", 0)]


@pytest.fixture(autouse=True)
def _reset_tile():
    yield
    ts.debug_set_tile(-1)


@pytest.mark.parametrize("tile", sorted(TILES))
@pytest.mark.parametrize("ci", range(len(CASES)))
@pytest.mark.parametrize("level", [0, 1, 2])
@pytest.mark.parametrize("form", range(len(FORMS)))
def test_tile_exact_on_integers(tile, ci, level, form):
    impl, flags = FORMS[form]
    if level == 0 and form > 0:
        pytest.skip("L0 runs one classical launch whatever the impl")
    ts.debug_set_tile(tile)
    case = cases.make_case(300 + ci, **CASES[ci])
    say(f"[tile {TILES[tile]}] case={ci} L={level} {impl} flags={flags}")
    a, b, c = inputs(case, 400 + ci, ints=True)
    got = run_gpu(case, a, b, c, 1.0, 1.0, level, impl, flags=flags)
    assert np.array_equal(got, einsum(case, a, b, c, 1.0, 1.0))


@pytest.mark.parametrize("tile", [2, 3])
@pytest.mark.parametrize("ci", range(len(CASES)))
@pytest.mark.parametrize("level", [0, 1, 2])
def test_rect_tile_float_vs_oracle(tile, ci, level):
    ts.debug_set_tile(tile)
    case = cases.make_case(320 + ci, **CASES[ci])
    say(f"[tile {TILES[tile]} f64] case={ci} L={level}")
    a, b, c = inputs(case, 500 + ci)
    for impl, flags in FORMS[:3]:
        got = run_gpu(case, a, b, c, 0.75, -0.5, level, impl, flags=flags)
        ref = run_oracle(case, a, b, c, 0.75, -0.5, level=level, impl=impl)
        assert cases.rel_err(got, ref) <= 1e-13, (impl, flags)


@pytest.mark.parametrize("tile", [2, 3])
@pytest.mark.parametrize("level", [0, 1, 2])
def test_rect_tile_fp32(tile, level):
    ts.debug_set_tile(tile)
    case = cases.make_case(340, m=(96, 4), n=(200,), k=(16, 10))
    a, b, c = inputs(case, 600)
    ref = einsum(case, a, b, c, 1.0, 1.0)
    for impl, flags in FORMS[:3]:
        got = run_gpu(case, a, b, c, 1.0, 1.0, level, impl, flags=flags, dtype=torch.float32)
        assert cases.rel_err(got, ref) <= 1e-5, (impl, flags)


def test_autotuner_offers_every_tile():
    case = cases.make_case(360, m=(64, 32), n=(64, 32), k=(64, 32))
    n = [int(np.prod(case[f"shape_{x}"])) for x in "ABC"]
    A, _ = dev(case["shape_A"], cases.uniform(n[0], 1))
    B, _ = dev(case["shape_B"], cases.uniform(n[1], 2))
    C, cbuf = dev(case["shape_C"], cases.uniform(n[2], 3))
    before = cbuf.clone()
    level, impl, tile = ts.autotune(A, case["idx_A"], B, case["idx_B"], C, case["idx_C"], measure=True)
    say(f"measured choice at 2048^3: L={level} impl={impl} tile={TILES[tile]}")
    assert tile in TILES
    assert torch.equal(cbuf, before)
    cands = ts.autotune_candidates(A, case["idx_A"], B, case["idx_B"], C, case["idx_C"])
    assert {t for (_, _, t, _) in cands} == set(TILES)
