"""Off-nominal ensemble (UQ) on the device against the oracle's ensemble.

FP64: every trajectory (nominal, exact samples through the element kernels
with their own moduli, approximate samples through the closure kernel) is
bit-identical to the oracle, so the bound is 0.  Mixed precision (tiled
assembly and element slots): u, v and f of every sample against the FP64
oracle within the mixed path's per-output bounds (relative to the field's
maximum over nodes).  On the device itself, approximate samples whose
coefficients select one force repeat that trajectory bit for bit in every
mode, and results do not depend on how the steps are split into calls.
"""

import numpy as np
import pytest

import nimblesm_amd as N
from nimblesm_amd.uq import UqModel
from oracle import oracle as O
from tests.conftest import load_golden
from tests.helpers import golden_oracle_model, max_rel_err, oracle_model_from_problem
from tests.test_gpu_model import gpu_model_from_golden

pytestmark = pytest.mark.gpu

FIELDS = ("displacement", "velocity", "internal_force")


def uq_wave_device(precision, steps, chunks=(None,)):
    g = load_golden("uq_wave")
    m = gpu_model_from_golden(g, precision)
    uq = UqModel()
    uq.ParseConfiguration(str(g["uq_model"]))
    uq.ParseBlockInput(str(g["uq_line"][0]), 1, {"density": float(g["rho"][0]), "bulk_modulus": float(g["K"][0]),
                                                 "shear_modulus": float(g["G"][0])})
    uq.Initialize(m, samples_text=str(g["uq_samples_text"]))
    integ = N.ExplicitTimeIntegrator(m, float(g["t_final"]), int(g["num_load_steps"]), int(g["output_frequency"]))
    integ.Initialize()
    done = 0
    for c in chunks:
        n = steps - done if c is None else c
        integ.Advance(n)
        done += n
    assert done == steps
    return g, uq, m


def uq_wave_oracle(g, uq, steps):
    K = [[uq.Parameters(1, s)["bulk_modulus"]] for s in range(uq.GetNumExactSamples())]
    G = [[uq.Parameters(1, s)["shear_modulus"]] for s in range(uq.GetNumExactSamples())]
    run = O.UqRun(golden_oracle_model(g), K, G, uq.interpolation_coefficients, float(g["t_final"]),
                  int(g["num_load_steps"]), int(g["output_frequency"]))
    return run.advance(steps)


def device_sample(m, s):
    return [m.GetUqNodeData(s, f) for f in FIELDS]


def nominal(m):
    return [m.GetVectorNodeData(f) for f in FIELDS]


def test_uq_wave_fp64_matches_oracle(hip_device):
    steps = 1500
    g, uq, m = uq_wave_device("fp64", steps, chunks=(700, 800))
    run = uq_wave_oracle(g, uq, steps)
    for got, want in zip(nominal(m), (run.nominal.u, run.nominal.v, run.nominal.f)):
        assert np.array_equal(got, want)
    for s in range(run.num_samples):
        for f, got, want in zip(FIELDS, device_sample(m, s), run.sample(s)):
            assert np.array_equal(got, want), (s, f, np.abs(got - want).max())
    # the nominal arrays still hold the nominal trajectory's F
    F = m.GetIntegrationPointData("deformation_gradient")
    Fo = golden_oracle_model(g).internal_force(run.nominal.u, want_ip=True)[1]
    assert np.array_equal(F, Fo)
    m.close()


def _check_selecting_samples(m, ne):
    """approximate samples with one-hot coefficients repeat a trajectory"""
    nom = nominal(m)
    for s, ref in ((ne + 0, device_sample(m, 0)), (ne + 2, nom), (ne + 4, device_sample(m, 1))):
        for f, got, want in zip(FIELDS, device_sample(m, s), ref):
            assert np.array_equal(got, want), (s, f)


@pytest.mark.parametrize("assembly", ["tiled", "slots"])
def test_uq_wave_mixed_matches_oracle(hip_device, monkeypatch, assembly):
    if assembly == "slots":
        monkeypatch.setenv("NSM_MIXED_ASSEMBLY", "slots")
    steps = 1500
    g, uq, m = uq_wave_device("mixed", steps)
    assert m.AssemblyInfo()["tiled"] == (assembly == "tiled")
    run = uq_wave_oracle(g, uq, steps)
    # mixed-path bounds on this deck: u 1e-6, v 1e-4, f 1e-4 of each field's max
    bounds = dict(displacement=1e-6, velocity=1e-4, internal_force=1e-4)
    for s in range(run.num_samples):
        for f, got, want in zip(FIELDS, device_sample(m, s), run.sample(s)):
            assert max_rel_err(got, want) <= bounds[f], (s, f, max_rel_err(got, want))
    _check_selecting_samples(m, uq.GetNumExactSamples())
    m.close()


def _cube_ensemble(precision, steps, chunks, seed=7):
    """2-block jittered cube, prescribed top displacement, two exact samples
    with other moduli per block and three approximate samples"""

    class Prob(N.CubeProblem):
        def bc_list(self):
            return [("prescribed_velocity", self.bottom, c, 0.0) for c in range(3)] + [
                ("prescribed_displacement", self.top, 2, 2e-4), ("initial_velocity", self.top, 0, 0.5)]

    prob = Prob(n=10, materials=[N.problems.STEEL, N.problems.ALUMINIUM], num_steps=steps, seed=seed)
    m = prob.make_model(precision)
    K0 = np.array([b.material.bulk_modulus for b in prob.blocks()])
    G0 = np.array([b.material.shear_modulus for b in prob.blocks()])
    K = np.stack([K0 * 1.3, K0 * 0.8])
    G = np.stack([G0 * 0.85, G0 * 1.1])
    c = np.array([[0.2, 0.5, 0.3], [-0.25, 1.0, 0.25], [0.0, 0.0, 1.0]])
    m.SetUqSamples(K, G, c)
    m.ComputeLumpedMass()
    dt = 0.5 * m.GetCriticalTimeStep()
    integ = N.ExplicitTimeIntegrator(m, steps * dt, steps, 7)
    integ.Initialize()
    for n in chunks:
        integ.Advance(n)
    om = oracle_model_from_problem(prob)
    run = O.UqRun(om, K, G, c, steps * dt, steps, 7, reaction_nodes=prob.top, reaction_coord=2)
    run.advance(steps)
    return m, run


def test_cube_ensemble_fp64_matches_oracle(hip_device):
    m, run = _cube_ensemble("fp64", 120, (50, 70))
    for got, want in zip(nominal(m), (run.nominal.u, run.nominal.v, run.nominal.f)):
        assert np.array_equal(got, want)
    for s in range(run.num_samples):
        for f, got, want in zip(FIELDS, device_sample(m, s), run.sample(s)):
            assert np.array_equal(got, want), (s, f, np.abs(got - want).max())
    # the reaction history belongs to the nominal trajectory
    R = m.GetReactionHistory(0, 120)
    assert max_rel_err(R, run.nominal.reaction) <= 1e-14
    m.close()


def test_cube_ensemble_mixed_matches_oracle_and_chunking(hip_device):
    m, run = _cube_ensemble("mixed", 120, (120,))
    m2, _ = _cube_ensemble("mixed", 120, (1, 59, 60))
    bounds = dict(displacement=1e-5, velocity=1e-4, internal_force=1e-4)
    for s in range(run.num_samples):
        got = device_sample(m, s)
        for f, a, want in zip(FIELDS, got, run.sample(s)):
            assert max_rel_err(a, want) <= bounds[f], (s, f, max_rel_err(a, want))
        for f, a, b in zip(FIELDS, got, device_sample(m2, s)):
            assert np.array_equal(a, b), (s, f)
    # third approximate sample selects exact sample 2
    for a, b in zip(device_sample(m, 4), device_sample(m, 1)):
        assert np.array_equal(a, b)
    m.close()
    m2.close()


def test_uq_deck_end_to_end(hip_device, tmp_path):
    """the reference's uq_wave deck (mesh, deck lines, samples file) through
    the deck reader, with its exact_* / sample_* output fields"""
    from scipy.io import netcdf_file

    from nimblesm_amd import deck as D
    from tests.test_gpu_deck import write_problem

    g = load_golden("uq_wave")
    fields = str(g["output_fields"]).split()
    path = write_problem(g, str(tmp_path), output_fields=fields)
    text = open(path).read().replace("number of load steps: 10000", "number of load steps: 2000")
    text = text.replace("final time: 1e-05", "final time: 2e-06")
    mat = f"mat_{int(g['blk_ids'][0])}"  # write_problem's material name for block 1
    text += f"uq parameters: {mat} {g['uq_line'][0]}\nuq model: {g['uq_model']}\n"
    open(path, "w").write(text)
    (tmp_path / str(g["uq_model"]).split()[1]).write_text(str(g["uq_samples_text"]))
    model, integ, out = D.run_deck(path, precision="fp64")
    assert integ.step == 2000 and model.uq.GetNumSamples() == 7
    assert (tmp_path / "parameter_samples.dat").exists()
    model.close()
    with netcdf_file(out, "r", mmap=False) as f:
        names = [D._chars_to_str(r) for r in f.variables["name_nod_var"][:]]
        vals = {n: f.variables[f"vals_nod_var{i + 1}"][:].copy() for i, n in enumerate(names)}
        T = len(f.variables["time_whole"][:])
    assert T == 22  # initial record, steps 0, 100, ..., 1900 and the last step
    assert {f"{q}_{c}" for q in fields for c in "xyz"} == set(names)
    for c in "xyz":
        # coefficient rows e_1, e_0, e_2
        assert np.array_equal(vals[f"sample_displacement_1_{c}"], vals[f"exact_displacement_1_{c}"])
        assert np.array_equal(vals[f"sample_displacement_3_{c}"], vals[f"displacement_{c}"])
        assert np.array_equal(vals[f"sample_force_3_{c}"], vals[f"internal_force_{c}"])
        assert np.array_equal(vals[f"sample_force_5_{c}"], vals[f"exact_force_2_{c}"])
    # the oracle's ensemble at the last record
    uq = UqModel()
    uq.ParseConfiguration(str(g["uq_model"]))
    uq.ParseBlockInput(str(g["uq_line"][0]), 1, {"bulk_modulus": float(g["K"][0]), "shear_modulus": float(g["G"][0])})
    uq.ReadSamples(str(g["uq_samples_text"]))
    uq.ScaleParameters()
    g2 = dict(g, num_load_steps=np.array(2000), t_final=np.array(2e-6))
    run = uq_wave_oracle(g2, uq, 2000)
    u2 = run.sample(3)[0]  # sample_displacement_2
    for i, c in enumerate("xyz"):
        assert np.array_equal(vals[f"sample_displacement_2_{c}"][-1], u2[:, i])
