"""Host logic of the ensemble workload and the control-variate estimator
(no device): the parameter draws follow the stated distributions, and the
estimator's mean, judged error and fallback are what their definitions give
on synthetic histories."""
import numpy as np

from nimblesm_amd.ensemble import Draws, config2_draws, draw_parameters, lognormal
from nimblesm_amd.surrogate import LinearSurrogate, Scaler, control_variate, features


def test_lognormal_mean_and_cov():
    x = lognormal(np.random.default_rng(0), 200e9, 0.1, 400_000)
    assert abs(x.mean() / 200e9 - 1) < 1e-3
    assert abs(x.std() / x.mean() - 0.1) < 2e-3


def test_draws_follow_the_workload():
    d = draw_parameters(20000, seed=1, n=(16, 24, 32, 40, 48))
    assert d.nu.min() >= 0.2 and d.nu.max() <= 0.45
    assert d.v_top.min() >= 0.5 and d.v_top.max() <= 2.0
    assert set(np.unique(d.n)) == {16, 24, 32, 40, 48}
    counts = np.bincount(d.n)[[16, 24, 32, 40, 48]]
    assert counts.min() > 0.18 * len(d) and counts.max() < 0.22 * len(d)
    assert abs(np.log(d.E / 200e9).std() - np.sqrt(np.log1p(0.01))) < 3e-3
    c2 = config2_draws()
    assert len(c2) == 320 and np.all(c2.n[:256] == 32)
    # seeded: the same draws every time
    assert np.array_equal(config2_draws().E, c2.E)


def test_features_standardise():
    d = draw_parameters(50000, seed=2, n=24)
    f = features(d)
    assert f.shape == (50000, 4)
    # log-features of mean-preserving lognormals are centred at -cov/2 = -0.05
    np.testing.assert_allclose(f.mean(axis=0), [-0.05, 0.0, -0.05, 0.0], atol=0.01)
    assert np.all(np.abs(f.std(axis=0) - 1) < 0.02)
    s = Scaler.fit(f)
    np.testing.assert_allclose(s.back(s.to(f)), f, rtol=1e-12, atol=1e-12)


class _Fixed:
    def __init__(self, pred):
        self.pred = pred

    def predict(self, d):
        return self.pred[: len(d)]


def synthetic(N=4000, seed=3):
    rng = np.random.default_rng(seed)
    d = draw_parameters(N, seed=seed, n=24)
    t = np.linspace(0, 1, 64)
    y = (d.v_top * np.sqrt(d.E * d.rho) * 1e-6)[:, None] * np.sin(3 * t)[None, :] * (1 + 0.3 * (d.nu - 0.3)[:, None])
    return d, y, rng


def test_control_variate_definition():
    d, y, rng = synthetic()
    f = y * (1 + 0.01 * rng.standard_normal(y.shape)) + 0.05 * y.mean()  # biased, noisy surrogate
    idx = np.sort(rng.choice(len(d), 200, replace=False))
    est = control_variate(_Fixed(f), d, lambda s: y[idx], plain_idx=idx, tol=1.0)
    r = y[idx] - f[idx]
    np.testing.assert_allclose(est.mean, f.mean(0) + r.mean(0), rtol=1e-14)
    se = r.std(0, ddof=1) / np.sqrt(200) * np.sqrt(1 - 200 / len(d))
    assert np.isclose(est.est_rel_err, np.linalg.norm(se) / np.linalg.norm(est.mean), rtol=1e-12)
    assert np.isclose(est.heldout_rel_err, np.linalg.norm(r) / np.linalg.norm(y[idx]), rtol=1e-12)
    assert not est.fallback and est.num_plain == 200
    # the bias is removed: far closer to the truth than the surrogate's own mean
    truth = y.mean(0)
    e_cv = np.linalg.norm(est.mean - truth) / np.linalg.norm(truth)
    e_f = np.linalg.norm(f.mean(0) - truth) / np.linalg.norm(truth)
    assert e_cv < 0.1 * e_f


def test_control_variate_falls_back():
    d, y, rng = synthetic()
    f = np.zeros_like(y)  # useless surrogate: large judged error
    idx = np.sort(rng.choice(len(d), 100, replace=False))
    calls = []

    def plain(s):
        calls.append(len(s))
        return y[idx] if len(s) == len(idx) else y

    est = control_variate(_Fixed(f), d, plain, plain_idx=idx, tol=1e-3)
    assert est.fallback and est.num_plain == len(d) and calls == [100, len(d)]
    np.testing.assert_array_equal(est.mean, y.mean(0))


def test_linear_surrogate_is_least_squares():
    d, y, _ = synthetic(500)
    lin = LinearSurrogate().fit(d, y)
    A = np.concatenate([np.ones((len(d), 1)), features(d)], axis=1)
    # residuals orthogonal to the design matrix
    assert np.abs(A.T @ (y - lin.predict(d))).max() < 1e-8 * np.abs(A.T @ y).max()


def test_draws_subset():
    d = draw_parameters(10, seed=4, n=(16, 24))
    s = d.subset(np.array([3, 1]))
    assert isinstance(s, Draws) and len(s) == 2 and s.E[0] == d.E[3] and s.n[1] == d.n[1]
