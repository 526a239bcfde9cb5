"""Checks of the oracle that follow from the mathematics alone (closed forms,
invariants, special cases), chosen so that a dropped term, a wrong sign or a
transposed operand fails at least one of them."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import UNIT_HEX, kf_to_mat, mat_to_kf, random_elements, sym_to_mat

K, G = 166.67e9, 76.92e9


def closed_form_stress(F, K, G):
    """sigma = K/2 (J - 1/J) I + G/J dev(J^{-2/3} F F^T) for one F (3x3)."""
    J = np.linalg.det(F)
    B = F @ F.T * J ** (-2.0 / 3.0)
    dev = B - np.trace(B) / 3.0 * np.eye(3)
    return 0.5 * K * (J - 1.0 / J) * np.eye(3) + G / J * dev


def test_gauss_tables():
    ip, N, dN = O.hex_tables()
    assert np.all(np.abs(np.abs(ip) - 0.577350269189626) == 0)
    assert np.allclose(N.sum(axis=1), 1.0, atol=1e-15)
    assert np.allclose(dN.sum(axis=1), 0.0, atol=1e-16)
    # N_a(xi_q) is the trilinear basis: interpolates node positions exactly
    xq = N @ UNIT_HEX
    assert np.allclose(xq, (ip + 1.0) / 2.0, atol=1e-15)


def test_pade_is_cos_acos_over_3():
    x = np.linspace(0.0, 1.0, 2001)
    approx = np.array([O.lib().orc_cos_of_acos_div3(t) for t in x])
    assert np.max(np.abs(approx - np.cos(np.arccos(x) / 3.0))) < 1e-15


def test_polar_decomposition_invariants():
    X, u = random_elements(400, seed=5)
    F = O.deformation_gradients(X, u).reshape(-1, 9)
    for p in range(0, len(F), 3):
        Fm = kf_to_mat(F[p])
        V, R = O.polar_decomp(F[p])
        Vm, Rm = sym_to_mat(V), kf_to_mat(R)
        assert np.linalg.norm(Vm @ Rm - Fm) < 1e-12  # CheckCorrectnessOfPolarDecomp tolerance
        assert np.allclose(Rm @ Rm.T, np.eye(3), atol=1e-12)
        assert np.all(np.linalg.eigvalsh(Vm) > 0)
        assert np.allclose(Vm @ Vm, Fm @ Fm.T, rtol=0, atol=1e-12 * np.linalg.norm(Fm) ** 2)


def test_polar_special_cases():
    V, R = O.polar_decomp(mat_to_kf(np.eye(3)))
    assert np.array_equal(V, [1, 1, 1, 0, 0, 0]) and np.array_equal(R, mat_to_kf(np.eye(3)))
    D = np.diag([1.2, 0.9, 1.05])
    V, _ = O.polar_decomp(mat_to_kf(D))
    assert np.allclose(sym_to_mat(V), D, atol=1e-15)
    th = 0.3
    Rz = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    V, R = O.polar_decomp(mat_to_kf(Rz @ D))  # F = R U  ->  V = R U R^T
    assert np.allclose(sym_to_mat(V), Rz @ D @ Rz.T, atol=1e-14)
    assert np.allclose(kf_to_mat(R), Rz, atol=1e-14)


def test_stress_closed_form():
    X, u = random_elements(300, seed=6, amps=(1e-3, 0.03))
    F = O.deformation_gradients(X, u).reshape(-1, 9)
    s = O.stress(K, G, F)
    for p in range(0, len(F), 5):
        ref = closed_form_stress(kf_to_mat(F[p]), K, G)
        assert np.linalg.norm(sym_to_mat(s[p]) - ref) <= 1e-11 * np.linalg.norm(ref)


def test_stress_special_cases():
    assert np.array_equal(O.stress(K, G, mat_to_kf(np.eye(3))), np.zeros(6))
    # pure rotation: stress-free
    th = 0.7
    Rz = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    assert np.max(np.abs(O.stress(K, G, mat_to_kf(Rz)))) < 1e-4  # ~1e-15 relative to K
    # uniform dilation lam: sigma = K/2 (J - 1/J) I, J = lam^3, no deviator
    lam = 1.01
    s = O.stress(K, G, mat_to_kf(lam * np.eye(3)))
    J = lam ** 3
    assert np.allclose(s[:3], 0.5 * K * (J - 1 / J), rtol=1e-14) and np.allclose(s[3:], 0, atol=1e-3)
    # simple shear gamma: sigma_xy = G * gamma / J^{5/3} with J = 1
    g = 1e-3
    Fs = np.eye(3)
    Fs[0, 1] = g
    s = O.stress(K, G, mat_to_kf(Fs))
    assert abs(s[3] - G * g) <= 1e-11 * G * g


def test_deformation_gradient_affine_exact():
    rng = np.random.default_rng(8)
    X, _ = random_elements(50, seed=9, amps=(0.0,))
    for e in range(len(X)):
        A = np.eye(3) + 0.05 * rng.normal(size=(3, 3))
        c = rng.normal(size=3)
        u = X[e] @ (A - np.eye(3)).T + c  # x = A X + c
        F = O.deformation_gradients(X[e:e + 1], u[None])[0]
        assert np.allclose(kf_to_mat(F), A[None], atol=1e-13)


def test_forces_translation_invariance_and_patch():
    X, u = random_elements(200, seed=10)
    F = O.deformation_gradients(X, u)
    s = O.stress(K, G, F)
    f = O.nodal_forces(X, u, s)
    scale = np.abs(f).max(axis=(1, 2))
    assert np.all(np.abs(f.sum(axis=1)).max(axis=1) <= 1e-12 * scale)  # sum_a f_a = 0
    # uniform stress on the unit cube: face load sigma_ij n_j * area/4 per node
    sig = np.tile(np.array([3.0, -2.0, 5.0, 0.5, -0.25, 1.5]), (1, 8, 1))
    f = O.nodal_forces(UNIT_HEX[None], np.zeros((1, 8, 3)), sig)[0]
    S = sym_to_mat(sig[0, 0])
    for a, Xa in enumerate(UNIT_HEX):
        n = np.where(Xa > 0.5, 1.0, -1.0)  # outward normals of the 3 faces at node a
        expect = -np.array([S[:, d] * n[d] for d in range(3)]).sum(axis=0) / 4.0
        assert np.allclose(f[a], expect, atol=1e-14)


def test_lumped_mass_total_and_regular_element():
    m = O.lumped_mass_elem(7800.0, UNIT_HEX * 0.5)
    assert np.allclose(m, 7800.0 * 0.125 / 8.0, rtol=1e-15)
    X, _ = random_elements(30, seed=12, amps=(0.0,))
    for e in range(len(X)):
        vol = 0.0
        # volume of a trilinear hex = sum over Gauss points of det J
        F = O.deformation_gradients(X[e:e + 1], np.zeros((1, 8, 3)))
        assert np.allclose(kf_to_mat(F[0]), np.eye(3), atol=1e-14)
        m = O.lumped_mass_elem(1.0, X[e])
        ip, N, dN = O.hex_tables()
        for q in range(8):
            J = X[e].T @ dN[q]
            vol += np.linalg.det(J)
        assert abs(m.sum() - vol) <= 1e-14 * vol


def test_characteristic_length():
    assert O.char_length(UNIT_HEX * 0.25) == 0.25
    X = UNIT_HEX * np.array([1.0, 2.0, 3.0])
    assert O.char_length(X) == 1.0


def test_model_level_invariants():
    X, conn = O.jittered_cube(4, 7, 0.1)
    M = O.Model(X, conn, [0, len(conn)], [7800.0], [K], [G])
    m = M.lumped_mass()
    ip, N, dN = O.hex_tables()
    assert abs(m.sum() - 7800.0) <= 1e-12 * 7800.0  # unit cube, total mass rho*V
    # rigid translation: forces vanish; affine stretch: interior nodes balance
    f = M.internal_force(np.tile([1e-3, -2e-3, 5e-4], (len(X), 1)))
    assert np.max(np.abs(f)) < 1e-3
    A = np.diag([1e-4, -5e-5, 2e-5])
    f = M.internal_force(X @ A.T)
    n = 4
    interior = [i + (n + 1) * (j + (n + 1) * k) for i in range(1, n) for j in range(1, n) for k in range(1, n)]
    assert np.max(np.abs(f[interior])) <= 1e-9 * np.max(np.abs(f))
    assert np.abs(f.sum(axis=0)).max() <= 1e-9 * np.abs(f).max()


def test_omp_path_is_bit_identical_to_serial():
    X, conn = O.jittered_cube(6, 3, 0.1)
    half = len(conn) // 2
    M = O.Model(X, conn, [0, half, len(conn)], [7800.0, 2700.0], [K, 68.63e9], [G, 26.32e9],
                np.arange(49, dtype=np.int32), np.full(49, 2, np.int32), np.ones(49, np.int32), np.full(49, 1.0))
    runs = []
    for threads in (1, 4):
        r = O.ExplicitRun(M, 50 * 1e-7, 50, 10, threads=threads, reaction_nodes=np.arange(49), reaction_coord=2)
        r.advance(50)
        runs.append(r)
    a, b = runs
    assert np.array_equal(a.u, b.u) and np.array_equal(a.v, b.v) and np.array_equal(a.f, b.f)
    assert np.array_equal(a.reaction, b.reaction)


def test_planted_mistakes_are_caught():
    """Each planted error in a copy of the stress fails the closed-form check."""
    X, u = random_elements(20, seed=13, amps=(1e-3,))
    F = O.deformation_gradients(X, u).reshape(-1, 9)[:40]
    s = O.stress(K, G, F)
    ref = np.array([closed_form_stress(kf_to_mat(Fp), K, G) for Fp in F])
    good = max(np.linalg.norm(sym_to_mat(s[p]) - ref[p]) / np.linalg.norm(ref[p]) for p in range(len(F)))
    assert good < 1e-11
    mistakes = {
        "sign of pressure": lambda S, Fp: S - 2 * np.trace(S) / 3 * np.eye(3),
        "transposed F": lambda S, Fp: closed_form_stress(Fp.T, K, G),
        "dropped J^-2/3": lambda S, Fp: S + G / np.linalg.det(Fp) * (1 - np.linalg.det(Fp) ** (-2 / 3))
        * (Fp @ Fp.T - np.trace(Fp @ Fp.T) / 3 * np.eye(3)),
        "K and G swapped": lambda S, Fp: closed_form_stress(Fp, G, K),
    }
    for name, plant in mistakes.items():
        bad = max(np.linalg.norm(plant(sym_to_mat(s[p]), kf_to_mat(F[p])) - ref[p]) / np.linalg.norm(ref[p])
                  for p in range(len(F)))
        assert bad > 1e-9, name


def _cbrt_libm_method(x):
    """Restatement of the cube-root method the FP64 device path uses (the
    platform libm's: polynomial start, one rational correction, 2^{k/3}
    factor, exact scaling) in float64 scalar arithmetic."""
    import math

    F = np.float64
    xm, xe = math.frexp(abs(float(x)))
    xm = F(xm)
    if x == 0 or not np.isfinite(x):
        return F(x) + F(x)
    u = (F(0.354895765043919860) + ((F(1.50819193781584896) + ((F(-2.11499494167371287) + ((F(2.44693122563534430) + (
        (F(-1.83469277483613086) + (F(0.784932344976639262) - F(0.145263899385486377) * xm) * xm) * xm)) * xm))
        * xm)) * xm))
    t2 = u * u * u
    cb2, sq2 = F(1.2599210498948731648), F(1.5874010519681994748)
    fac = [F(1.0) / sq2, F(1.0) / cb2, F(1.0), cb2, sq2][2 + int(math.fmod(xe, 3))]
    ym = u * (t2 + F(2.0) * xm) / (F(2.0) * t2 + xm) * fac
    q = int(xe / 3) if xe >= 0 else -int(-xe / 3)
    return F(math.ldexp(float(ym if x > 0 else -ym), q))


def test_cbrt_method_matches_platform_libm():
    """The stress calls std::cbrt (src/nimble_material.cc:274); the device FP64
    path reproduces the x86-64 GNU libm method bit for bit.  That libm result is
    not correctly rounded, so a correctly rounded cube root would not do."""
    import ctypes

    libm = ctypes.CDLL("libm.so.6")
    libm.cbrt.restype = ctypes.c_double
    libm.cbrt.argtypes = [ctypes.c_double]
    rng = np.random.default_rng(17)
    xs = np.concatenate([1 + rng.uniform(-1e-3, 1e-3, 20000), 1 + rng.uniform(-1e-9, 1e-9, 5000),
                         np.exp(rng.uniform(-600, 600, 10000)), -rng.uniform(0.1, 10, 5000), [1.0, 8.0, 0.125]])
    got = np.array([_cbrt_libm_method(x) for x in xs])
    ref = np.array([libm.cbrt(float(x)) for x in xs])
    assert np.array_equal(got, ref)
    # and the method is not the correctly rounded cube root
    assert np.sum(got[:20000] != np.cbrt(xs[:20000])) > 1000


# ------------------------------------------------------------- reaction sum
def affine_cube(n, F, materials):
    """Undistorted n^3 unit cube (no jitter) with blocks in equal z-layers,
    displacement u = (F - I) X, no conditions: the first step leaves u as it
    is (v = a = 0), so the step's internal force is that of x = F X."""
    X, conn = O.jittered_cube(n, 1, 0.0)
    nb = len(materials)
    per = n * n
    off = [(n * b // nb) * per for b in range(nb + 1)]
    rho = [m[0] for m in materials]
    Kb = [m[1] for m in materials]
    Gb = [m[2] for m in materials]
    M = O.Model(X, conn, off, rho, Kb, Gb)
    u = X @ (F - np.eye(3)).T
    n1 = n + 1
    layer = lambda k: np.arange(n1 * n1, dtype=np.int32) + k * n1 * n1
    return M, u, layer(n), layer(0)


def oracle_reaction(M, u, nodes, coord):
    """R = -sum_{nodes} f_int[coord] of the oracle's first step from x = F X."""
    r = O.ExplicitRun(M, 1e-9, 1, 0, reaction_nodes=nodes, reaction_coord=coord)
    r.u[:] = u
    r.advance(1)
    assert np.array_equal(r.u, u)  # the state the force was evaluated at
    return r.reaction[0]


def traction_resultant(F, Kb, Gb):
    S = sym_to_mat(O.stress(Kb, Gb, mat_to_kf(F)))
    cof = np.linalg.det(F) * np.linalg.inv(F).T
    return S @ cof @ np.array([0.0, 0.0, 1.0])


def closed_form_reaction(F, Kb, Gb, axis):
    """Traction resultant on the face Z = 1 of the unit cube under x = F X
    with the oracle's (uniform) stress: Nanson's n da = J F^{-T} N dA with
    N = e_Z, A = 1, so R = e_axis . sigma . J F^{-T} e_Z.  This equals
    -sum_{top nodes} f_int,axis exactly for an undistorted mesh: the nodal
    sum of B^T sigma is sigma . int grad(phi) dv with phi the sum of the
    top nodes' shape functions (1 on Z = 1, 0 below the top element layer),
    which 2x2x2 Gauss integration evaluates exactly."""
    return traction_resultant(F, Kb, Gb)[axis]


AFFINE = {
    "uniaxial": np.diag([1.0, 1.0, 1.0015]),
    "simple shear": np.array([[1.0, 0.0, 0.02], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]]),
    "general": np.array([[1.0010, 0.0003, 0.0002], [-0.0001, 0.9995, 0.0004], [0.0002, -0.0003, 1.0020]]),
}
STEEL_KG = (7800.0, 166.67e9, 76.92e9)
ALU_KG = (2700.0, 68.63e9, 26.32e9)


@pytest.mark.parametrize("case", sorted(AFFINE))
def test_reaction_sum_closed_form(case):
    """The reaction history's sum (R = -sum_top f_int,z, a quantity the
    reference does not output) against the traction resultant of a uniform
    affine state: R_top = e_z . sigma(F) . J F^{-T} e_Z, and R_bottom =
    -R_top for one material, to 1e-12 of the traction resultant's magnitude
    (roundoff scales with the whole stress: under simple shear R_z =
    -G gamma^2 / 3 is 1/150 of the shear traction)."""
    F = AFFINE[case]
    M, u, top, bottom = affine_cube(4, F, [STEEL_KG])
    scale = np.linalg.norm(traction_resultant(F, STEEL_KG[1], STEEL_KG[2]))
    for axis in (2, 0):
        R = oracle_reaction(M, u, top, axis)
        ref = closed_form_reaction(F, STEEL_KG[1], STEEL_KG[2], axis)
        assert abs(R - ref) <= 1e-12 * scale, (axis, R, ref)
        Rb = oracle_reaction(M, u, bottom, axis)
        assert abs(Rb + R) <= 1e-12 * scale, (axis, Rb, R)
    assert abs(closed_form_reaction(F, STEEL_KG[1], STEEL_KG[2], 2)) > 1e-4 * scale


def test_reaction_sum_planted_mistakes_fail():
    """A sign flip, the bottom face instead of the top, the x axis instead of
    z, and the lower block's material all miss the closed form of the top
    face by far more than its tolerance (two blocks: steel below, aluminium
    on top; general affine F)."""
    F = AFFINE["general"]
    M, u, top, bottom = affine_cube(4, F, [STEEL_KG, ALU_KG])
    ref = closed_form_reaction(F, ALU_KG[1], ALU_KG[2], 2)
    good = oracle_reaction(M, u, top, 2)
    assert abs(good - ref) <= 1e-12 * abs(ref)
    planted = {
        "sign flip": -good,
        "bottom face": oracle_reaction(M, u, bottom, 2),
        "x axis": oracle_reaction(M, u, top, 0),
        "lower block's material": closed_form_reaction(F, STEEL_KG[1], STEEL_KG[2], 2),
    }
    for name, val in planted.items():
        assert abs(val - ref) > 1e-3 * abs(ref), name
