"""CPU: the sampled FP64 check used by the full-size GPU tests equals numpy.einsum on small cases."""
import numpy as np
import pytest

import cases
import sampled


@pytest.mark.parametrize("groups", [dict(m=(6, 5), n=(7,), k=(4, 3)), dict(m=(9,), n=(4, 5), k=(2, 3, 4)),
                                    dict(m=(16, 16, 32), n=(3,), k=(5,))])
def test_sampled_values_equal_einsum(groups):
    case = cases.make_case(31, **groups)
    n = [int(np.prod(case[f"shape_{x}"])) for x in "ABC"]
    a, b = cases.uniform(n[0], 1), cases.uniform(n[1], 2)
    want = cases.einsum_ref(case, cases.host_tensor(case["shape_A"], a), cases.host_tensor(case["shape_B"], b),
                            cases.host_tensor(case["shape_C"], np.zeros(n[2])), 1.0, 0.0)
    flat = np.asarray(want).ravel(order="F")
    off, ref = sampled.sample(case, a, b, nsamp=200, seed=3, chunk=64)
    assert np.allclose(ref, flat[off], rtol=1e-14, atol=1e-14)
    # a planted mistake (B's K vector shifted by one position) is seen
    (m, nn, k), (am, ak, bk, bn, cm, cn) = sampled.describe_case(case)
    bad = np.array([a[am[i] + ak] @ b[np.roll(bk, 1) + bn[j]] for i, j in zip(range(3), range(3))])
    good = np.array([flat[cm[i] + cn[j]] for i, j in zip(range(3), range(3))])
    assert not np.allclose(bad, good)
