"""Off-nominal ensemble (UQ): host-side samples file / parameter scaling /
deck keys, and the oracle's ensemble (CPU only).

The reference's UQ regression deck (test/uq/uq_wave: mesh, deck and samples
file in tests/golden/uq_wave.npz) has no gold results in the reference
checkout, so the oracle's ensemble is pinned by what the closure fixes
exactly.  An approximate sample whose coefficients select one force
(c = e_0 or e_{k+1}) receives that force bit for bit (0 * x = +-0, 1 * x = x),
and it shares the mass, the kicks and the velocity conditions, so it repeats
the nominal or exact-sample trajectory bit for bit.  A wrong coefficient
index, a dropped term or a different update order fails these.  Exact samples
are nominal runs with other moduli (the oracle's nominal run is pinned
against the reference's gold files elsewhere).  That prescribed
displacements leave the samples' velocities alone is the reference's
ApplyKinematicBC (src/nimble_boundary_condition_manager.h:184-215); it is
checked here only as the observable difference it makes.
"""

import numpy as np
import pytest

from nimblesm_amd import deck as D
from nimblesm_amd.uq import UqModel, block_id_of_material
from oracle import oracle as O
from tests.conftest import load_golden
from tests.helpers import golden_oracle_model

UQ_WAVE_K = 1.0e12


def uq_wave_model():
    g = load_golden("uq_wave")
    uq = UqModel()
    uq.ParseConfiguration(str(g["uq_model"]))
    assert str(g["uq_material"][0]) == "material_1"
    uq.ParseBlockInput(str(g["uq_line"][0]), 1, {"density": float(g["rho"][0]), "bulk_modulus": float(g["K"][0]),
                                                 "shear_modulus": float(g["G"][0])})
    uq.ReadSamples(str(g["uq_samples_text"]))
    uq.ScaleParameters()
    return g, uq


def test_uq_wave_samples_file():
    g, uq = uq_wave_model()
    assert uq.samples_file_name == "uq_wave.coef"
    assert uq.parameter_order == [(1, "bulk_modulus")]
    assert uq.parameter_uncertainties == {1: {"bulk_modulus": 5.0e11}}
    assert (uq.GetNumExactSamples(), uq.napprox_samples, uq.GetNumSamples()) == (2, 5, 7)
    assert uq.parameter_samples == [[-1.0], [1.0], [-1.0], [-0.5], [0.0], [0.5], [1.0]]
    assert uq.interpolation_coefficients == [[0.0, 1.0, 0.0], [0.5, 0.5, 0.0], [1.0, 0.0, 0.0], [0.5, 0.0, 0.5],
                                             [0.0, 0.0, 1.0]]
    # p = p0 + range * xi, other parameters untouched
    K = [uq.Parameters(1, s)["bulk_modulus"] for s in range(7)]
    assert K == [5.0e11, 1.5e12, 5.0e11, 7.5e11, 1.0e12, 1.25e12, 1.5e12]
    assert all(uq.Parameters(1, s)["shear_modulus"] == float(g["G"][0]) for s in range(7))
    assert uq.nominal_parameters[1]["bulk_modulus"] == UQ_WAVE_K


def test_write_samples_listing(tmp_path):
    _, uq = uq_wave_model()
    path = tmp_path / "parameter_samples.dat"
    uq.WriteSamples(str(path))
    lines = path.read_text().splitlines()
    assert lines[0] == "# bulk_modulus_1 "
    assert lines[1:] == ["   5e+11 ", " 1.5e+12 ", "   5e+11 ", " 7.5e+11 ", "   1e+12 ", "1.25e+12 ", " 1.5e+12 "]


@pytest.mark.parametrize("text, msg", [
    ("1 1\n0.5 0.2\n0.0 : 1.0 0.0\n", "not enough exact sample parameters on line 2 expecting 1 found 2"),
    ("1 1\n0.5\n0.0 1.0 0.0\n", "wrong format (no colon) on line 3"),
    ("1 1\n0.5\n0.0 0.1 : 1.0 0.0\n", "has 2 parameters (preceding :) on line 3 expecting 1"),
    ("1 1\n0.5\n0.0 : 1.0\n", "has 1 coefficients on line 3 expecting 2"),
    ("1 2\n0.5\n0.0 : 1.0 0.0\n", "ends before line 4"),
])
def test_samples_file_errors(text, msg):
    uq = UqModel()
    uq.ParseConfiguration("file s.coef")
    uq.ParseBlockInput("bulk_modulus 1.0", 1, {"bulk_modulus": 2.0, "shear_modulus": 1.0})
    with pytest.raises(ValueError, match=msg.replace("(", r"\(").replace(")", r"\)")):
        uq.ReadSamples(text)


def test_block_input_errors():
    uq = UqModel()
    with pytest.raises(ValueError):
        uq.ParseConfiguration("sobol 10")
    with pytest.raises(ValueError, match="no uq parameters"):
        uq.ParseBlockInput("none", 3, {})
    with pytest.raises(ValueError, match="pairs"):
        uq.ParseBlockInput("bulk_modulus", 3, {})


def test_block_id_of_material():
    blocks = [(3, "steel"), (1, "alu"), (2, "steel")]
    assert block_id_of_material(blocks, "steel") == 2  # first in ascending block id
    assert block_id_of_material(blocks, "alu") == 1
    assert block_id_of_material(blocks, "wood") == -1


def test_deck_reads_uq_keys(tmp_path):
    text = "\n".join([
        "genesis input file: mesh.g", "number of load steps: 10",
        "material parameters: material_1 neohookean density 7.8 shear_modulus 1.5e12 bulk_modulus 1.0e12",
        "uq parameters:              material_1 bulk_modulus 5.0e11 ",
        "uq model:                   file uq_wave.coef",
        "element block: block_1 material_1", ""])
    p = tmp_path / "d.in"
    p.write_text(text)
    deck = D.parse_input_deck(str(p))
    assert deck.uq_model == "file uq_wave.coef"
    assert deck.uq_parameters == {"material_1": "bulk_modulus 5.0e11"}
    p.write_text(text + "contact: on\n")
    with pytest.raises(ValueError, match="outside this path"):
        D.parse_input_deck(str(p))


def _uq_wave_oracle(steps, coeff=None, K_exact=None):
    g, uq = uq_wave_model()
    model = golden_oracle_model(g)
    K = [[uq.Parameters(1, s)["bulk_modulus"]] for s in range(2)] if K_exact is None else K_exact
    G = [[float(g["G"][0])]] * len(K)
    c = uq.interpolation_coefficients if coeff is None else coeff
    run = O.UqRun(model, K, G, c, float(g["t_final"]), int(g["num_load_steps"]), int(g["output_frequency"]))
    run.advance(steps)
    return run


def test_oracle_closure_selects_trajectories():
    run = _uq_wave_oracle(1200)
    nom, (e1, e2) = run.nominal, run.exact
    pairs = {0: e1, 2: nom, 4: e2}  # coefficients e_1, e_0, e_2
    for s, r in pairs.items():
        for a, b in zip(run.sample(2 + s), (r.u, r.v, r.f)):
            assert np.array_equal(a, b), s
    # the interpolating samples differ from all three and move
    for s in (1, 3):
        u = run.sample(2 + s)[0]
        assert not np.array_equal(u, nom.u) and np.abs(u).max() > 0
    # exact samples with other bulk moduli differ from the nominal run
    assert not np.array_equal(e1.u, nom.u) and not np.array_equal(e2.u, nom.u)
    assert run.time_current == nom.time


def test_oracle_exact_sample_with_nominal_moduli_is_nominal():
    run = _uq_wave_oracle(600, coeff=[[0.25, 0.75]], K_exact=[[UQ_WAVE_K]])
    assert np.array_equal(run.exact[0].u, run.nominal.u)
    assert np.array_equal(run.exact[0].f, run.nominal.f)
    assert np.array_equal(run.u[0], run.nominal.u)


def test_oracle_samples_ignore_prescribed_displacement():
    import nimblesm_amd as N
    from tests.helpers import oracle_model_from_problem

    class Prob(N.CubeProblem):
        def bc_list(self):
            return [("prescribed_velocity", self.bottom, c, 0.0) for c in range(3)] + [
                ("prescribed_displacement", self.top, 2, 1e-4)]

    prob = Prob(n=3, materials=[N.problems.STEEL], num_steps=40)
    om = oracle_model_from_problem(prob)
    dt = 0.5 * om.critical_dt()
    # output every step: the end-of-step velocity conditions apply too
    run = O.UqRun(om, np.zeros((0, 1)), np.zeros((0, 1)), [[1.0]], 40 * dt, 40, 1)
    run.advance(40)
    top, bottom = prob.top, prob.bottom
    # the nominal top face follows the prescribed displacement, the sample's
    # (same forces, no displacement condition) does not
    assert np.allclose(run.nominal.u[top, 2], 1e-4, rtol=1e-12, atol=0)
    assert np.abs(run.u[0][top, 2] - 1e-4).max() > 1e-6
    # velocity conditions do apply to samples
    assert np.all(run.v[0][bottom] == 0.0) and np.all(run.u[0][bottom] == 0.0)
