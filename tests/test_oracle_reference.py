"""The oracle against NimbleSM's own element / material code (oracle/_ref),
bit for bit.  Skipped where the reference tree was not available to build
oracle/_ref."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import kf_to_mat, random_elements

# Only checks whether oracle/_ref was built: the library itself is loaded by
# the fixture, so collecting this module (e.g. for a GPU-only run that
# deselects it) maps no reference-built code.
pytestmark = pytest.mark.skipif(not os.path.exists(O.REF_PATH),
                                reason="oracle/_ref not built (reference sources absent)")


@pytest.fixture(scope="module")
def R():
    return O.ref_lib()


@pytest.fixture(scope="module")
def elems():
    return random_elements(3000, seed=11)


def test_deformation_gradients_bit_exact(R, elems):
    X, u = elems
    F = O.deformation_gradients(X, u)
    Fr = np.zeros_like(F)
    R.ref_deformation_gradients(len(X), X, u, Fr)
    assert np.array_equal(F, Fr)


def test_stress_bit_exact(R, elems):
    X, u = elems
    F = O.deformation_gradients(X, u)
    for K, G in [(1.6e11, 0.8e11), (166.67e9, 76.92e9), (10.0e9, 1.034e9)]:
        s = O.stress(K, G, F)
        sr = np.zeros_like(s)
        R.ref_stress(F.size // 9, K, G, F, sr)
        assert np.array_equal(s, sr)


def test_polar_decomp_bit_exact(R, elems):
    X, u = elems
    F = O.deformation_gradients(X, u).reshape(-1, 9)[:5000]
    V, Rm = np.zeros((len(F), 6)), np.zeros((len(F), 9))
    R.ref_polar_decomp(len(F), np.ascontiguousarray(F), V, Rm)
    for p in range(0, len(F), 7):
        v, r = O.polar_decomp(F[p])
        assert np.array_equal(v, V[p]) and np.array_equal(r, Rm[p])


def test_nodal_forces_bit_exact(R, elems):
    X, u = elems
    s = O.stress(1.6e11, 0.8e11, O.deformation_gradients(X, u))
    f = O.nodal_forces(X, u, s)
    fr = np.zeros_like(f)
    R.ref_nodal_forces(len(X), X, u, s, fr)
    assert np.array_equal(f, fr)


def test_lumped_mass_and_char_length_bit_exact(R, elems):
    X, u = elems
    E = 500
    m = np.array([O.lumped_mass_elem(7800.0, X[e]) for e in range(E)])
    mr = np.zeros((E, 8))
    R.ref_lumped_mass(E, 7800.0, np.ascontiguousarray(X[:E]), mr)
    assert np.array_equal(m, mr)
    x = np.ascontiguousarray(X + u)
    cl = np.array([O.char_length(x[e]) for e in range(len(x))])
    clr = np.zeros(len(x))
    R.ref_char_length(len(x), x, clr)
    assert np.array_equal(cl, clr)


def test_scalar_kernels_bit_exact(R):
    rng = np.random.default_rng(3)
    for x in np.concatenate([np.linspace(0, 1, 257), rng.uniform(0, 1, 200)]):
        assert O.lib().orc_cos_of_acos_div3(x) == R.ref_cos_of_acos_div3(x)
    for _ in range(200):
        m = 0.5 * rng.normal(size=9) + 2.0 * np.eye(3).reshape(9)
        # both interpretations (row-major and K_F storage) must have det > 0:
        # the reference asserts on a non-positive determinant
        if np.linalg.det(m.reshape(3, 3)) <= 0.0 or np.linalg.det(kf_to_mat(m)) <= 0.0:
            continue
        inv_o = np.zeros(9)
        inv_r = np.zeros(9)
        assert O.lib().orc_invert_full33(m, inv_o) == R.ref_invert_full33(m, inv_r)
        assert np.array_equal(inv_o, inv_r)
        d, inv3, st = O.invert3x3(m)
        inv3r = np.zeros(9)
        assert d == R.ref_invert3x3(m, inv3r)
        assert np.array_equal(inv3.reshape(9), inv3r)
