"""The 64-point quantity of interest: the reaction history sampled at
equally spaced times (sample k at (k+1) n / 64 steps, linear between steps,
0 at rest before step 0).  The rule is defined with the device sampler (the
reference has no reaction output); the oracle's restatement is pinned here by
closed forms, the GPU sampler against it in tests/test_gpu_model.py."""
import numpy as np
import pytest

from oracle import oracle as O


def sampled_linear(a, n, m):
    return np.array([a * (k + 1) * n / m for k in range(m)])


@pytest.mark.parametrize("n,m", [(400, 64), (40, 64), (7, 3), (64, 64), (1, 5)])
def test_linear_history_is_reproduced_exactly(n, m):
    # value a*j at integer step j, i.e. R[j-1] = a*j: interpolation is exact
    a = 3.25
    R = a * np.arange(1, n + 1, dtype=np.float64)
    np.testing.assert_allclose(O.sample_history(R, m), sampled_linear(a, n, m), rtol=1e-15, atol=0)


def test_stride_of_400_steps_in_64_samples():
    # 400 / 64 = 6.25 steps per sample: sample 0 sits between steps 6 and 7
    rng = np.random.default_rng(0)
    R = rng.normal(size=400)
    s = O.sample_history(R, 64)
    assert s[0] == 0.75 * R[5] + 0.25 * R[6]
    assert s[3] == R[24]  # p = 25.0 exactly: value after step 24
    assert s[-1] == R[-1]  # the last sample is the end of the run


def test_first_samples_start_from_rest():
    # fewer steps than samples: early samples interpolate from the rest value 0
    s = O.sample_history(np.array([2.0, 2.0]), 4)  # p = 0.5, 1, 1.5, 2
    assert np.array_equal(s, [1.0, 2.0, 2.0, 2.0])


def test_planted_off_by_one_is_caught():
    # a sampler that reads R[j] at integer step j (forgetting the rest value)
    def wrong(R, m):
        n = len(R)
        out = []
        for k in range(m):
            p = (k + 1) * n / m
            j = min(int(p), n - 1)
            w = p - j
            out.append((1 - w) * R[j] + w * R[min(j + 1, n - 1)])
        return np.array(out)

    n, m, a = 400, 64, 1.5
    R = a * np.arange(1, n + 1, dtype=np.float64)
    assert not np.allclose(wrong(R, m), sampled_linear(a, n, m), rtol=1e-12)
