"""The learned control variate on the 24^3 cube workload: a network trained
on the device on 512 plain FP64 solves maps (E, nu, rho, v_top) to the
64-point reaction history; 8192 draws are estimated with 5% plain solves for
bias correction and error judging.

Truth is the mean history of plain solves of all 8192 draws (batched, mixed
precision, which is within 1e-7 of FP64 per sample: test_gpu_ensemble).  The
estimate must be within 1e-3 of it, its judged error must be below 1e-3, and
at equal plain-solve cost it must beat plain Monte Carlo and the linear
control variate (the linear map in place of the network); a tolerance it
cannot meet makes it fall back to plain solves of every draw.
"""
import numpy as np
import pytest

from nimblesm_amd.ensemble import EnsembleProblem, draw_parameters
from nimblesm_amd.surrogate import LinearSurrogate, MLPSurrogate, control_variate
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu

STEPS = 400


def solve(d, precision):
    prob = EnsembleProblem(d, num_steps=STEPS, precision=precision, tile_elems=512).build()
    q = prob.run()
    prob.close()
    return q


def test_control_variate_on_the_24_cube(hip_device):
    train = draw_parameters(512, seed=101, n=24)
    ev = draw_parameters(8192, seed=202, n=24)
    y_train = solve(train, "fp64")
    truth_all = solve(ev, "mixed")
    truth = truth_all.mean(axis=0)
    plain = lambda d: solve(d, "mixed")  # noqa: E731

    mlp = MLPSurrogate(seed=0).fit(train, y_train, iterations=3000)
    assert mlp.loss[-1] < 1e-3 * mlp.loss[0]
    pred = mlp.predict(ev)
    idx = np.sort(np.random.default_rng(7).choice(len(ev), size=410, replace=False))
    # the plain solves of the correction subset: the same trajectories as in the truth
    est = control_variate(mlp, ev, lambda d: truth_all[idx], pred=pred, plain_idx=idx)
    err_mlp = rel_err(est.mean, truth)
    # equal cost: 512 + 410 plain solves
    lin = LinearSurrogate().fit(train, y_train)
    est_lin = control_variate(lin, ev, lambda d: truth_all[idx], plain_idx=idx)
    err_lin = rel_err(est_lin.mean, truth)
    mc = np.random.default_rng(8).choice(len(ev), size=512 + 410, replace=False)
    err_mc = rel_err(truth_all[mc].mean(axis=0), truth)
    surrogate_only = rel_err(pred.mean(axis=0), truth)
    print(f"mean-history error: network CV {err_mlp:.2e} (judged {est.est_rel_err:.2e}, held-out network error "
          f"{est.heldout_rel_err:.2e}), linear CV {err_lin:.2e}, plain MC {err_mc:.2e}, network alone "
          f"{surrogate_only:.2e}", flush=True)
    assert not est.fallback and est.num_plain == 410
    assert err_mlp <= 1e-3 and est.est_rel_err <= 1e-3
    assert err_mlp * 5 < err_mc and err_mlp < err_lin
    # the plain-solve callback really solves: two draws against the truth run
    two = ev.subset(idx[:2])
    assert np.array_equal(plain(two), truth_all[idx[:2]])
    # a tolerance the estimate cannot meet falls back to plain solves of every draw
    fb = control_variate(mlp, ev, lambda d: truth_all[idx] if len(d) == len(idx) else truth_all, pred=pred,
                         plain_idx=idx, tol=1e-9)
    assert fb.fallback and fb.num_plain == len(ev) and np.array_equal(fb.mean, truth)
    mlp.close()
