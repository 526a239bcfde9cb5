"""The allocation-free reaction sampling entry point against the allocating
one, through the C ABI: bit-identical results on two calls in a row, the
second with a smaller sample count (the fixed device array keeps the first
call's values past the second's count, which must not leak into it), and the
size limit rejected with NSM_ERR_ARG."""
import ctypes as C

import numpy as np
import pytest

import nimblesm_amd as N
from nimblesm_amd import _capi

pytestmark = pytest.mark.gpu


def _sample(fn, model, steps, count):
    out = np.full(count, np.nan)
    st = fn(model._h, steps, count, out.ctypes.data)
    assert st == _capi.NSM_OK, _capi.lib().nsm_last_error().decode()
    return out


def test_noalloc_sampling_matches_allocating_entry_point(hip_device):
    lib = N.lib()
    for precision in ("fp64", "mixed"):
        prob = N.CubeProblem(n=6, materials=[N.problems.STEEL], num_steps=40, name="noalloc")
        m = prob.make_model(precision)
        prob.make_integrator(m).Integrate()
        for steps, count in ((40, 64), (23, 7)):
            ref = _sample(lib.nsm_model_sample_reaction_history, m, steps, count)
            got = _sample(lib.nsm_model_sample_reaction_history_noalloc, m, steps, count)
            assert np.isfinite(ref).all() and np.any(ref != 0.0)
            assert np.array_equal(got, ref), (precision, steps, count)
        # the limit and the argument checks
        big = np.zeros(_capi.NSM_NOALLOC_MAX_SAMPLES + 1)
        st = lib.nsm_model_sample_reaction_history_noalloc(m._h, 40, len(big), big.ctypes.data)
        assert st == _capi.NSM_ERR_ARG
        assert b"allocation-free maximum" in lib.nsm_last_error()
        st = lib.nsm_model_sample_reaction_history_noalloc(m._h, 41, 4, big.ctypes.data)
        assert st == _capi.NSM_ERR_ARG
        # a full-size call stays in bounds and agrees
        n = _capi.NSM_NOALLOC_MAX_SAMPLES
        ref = _sample(lib.nsm_model_sample_reaction_history, m, 40, n)
        got = _sample(lib.nsm_model_sample_reaction_history_noalloc, m, 40, n)
        assert np.array_equal(got, ref)
        m.close()
