"""Autotuner on the GPU: a measured choice is one of the candidates, is cached per shape, and calls
with level=-1 / impl="auto" run it and stay within the FP64 tolerance of the classical result."""
import numpy as np
import pytest

import cases
import tblis_strassen as ts

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

from test_gpu_parity import dev  # noqa: E402


@pytest.mark.parametrize("groups", [dict(m=(32, 32), n=(32, 32), k=(32, 32)), dict(m=(24, 40), n=(56,), k=(30, 20))])
def test_measured_choice_runs_and_is_accurate(groups):
    case = cases.make_case(77, **groups)
    n = [int(np.prod(case[f"shape_{x}"])) for x in "ABC"]
    A, _ = dev(case["shape_A"], cases.uniform(n[0], 1))
    B, _ = dev(case["shape_B"], cases.uniform(n[1], 2))
    C, cbuf = dev(case["shape_C"], cases.uniform(n[2], 3))
    before = cbuf.clone()
    level, impl, tile = ts.autotune(A, case["idx_A"], B, case["idx_B"], C, case["idx_C"], measure=True)
    print("measured choice", level, impl, tile, flush=True)
    assert level in (0, 1, 2) and impl in (0, 3, 4) and tile in (0, 1)
    assert torch.equal(cbuf, before)  # tuning ran on a scratch copy
    assert ts.autotune(A, case["idx_A"], B, case["idx_B"], C, case["idx_C"]) == (level, impl, tile)  # cached
    ref, rbuf = dev(case["shape_C"], cases.uniform(n[2], 3))
    ts.tensor_mult(1.0, A, case["idx_A"], B, case["idx_B"], 1.0, ref, case["idx_C"], level=0, flags=ts.FLAG_SYNC)
    ts.tensor_mult(1.0, A, case["idx_A"], B, case["idx_B"], 1.0, C, case["idx_C"], level=-1, impl="auto",
                   flags=ts.FLAG_SYNC)
    err = float(torch.linalg.norm(cbuf - rbuf) / torch.linalg.norm(rbuf))
    assert err <= 5e-15 * 2 ** level
