"""Device ModelData + explicit driver against the oracle.

FP64: nodal forces, displacements, velocities, F and stress are bit-identical
to the oracle (the reference's FP64 run on the same input); the reaction
history, a parallel tree sum on the device, agrees to 1e-14 relative.  Mixed precision (FP32 element math, FP64 node state): errors
against the FP64 oracle are bounded per output as stated in each test.
"""
import numpy as np
import pytest

import nimblesm_amd as N
from oracle import oracle as O
from tests.conftest import load_golden
from tests.helpers import golden_oracle_model, max_rel_err, oracle_model_from_problem, output_steps, rel_err

pytestmark = pytest.mark.gpu

KINDS = {0: "initial_velocity", 1: "prescribed_velocity", 2: "prescribed_displacement"}


def gpu_model_from_golden(g, precision="fp64"):
    blocks = []
    for b in range(len(g["rho"])):
        e0, e1 = int(g["blk_off"][b]), int(g["blk_off"][b + 1])
        blocks.append(N.Block(int(g["blk_ids"][b]), g["conn"][e0:e1],
                              N.NeohookeanMaterial(g["rho"][b], g["K"][b], g["G"][b])))
    m = N.ModelData(g["X"], blocks, precision)
    for node, c, k, v in zip(g["bc_node"], g["bc_coord"], g["bc_kind"], g["bc_value"]):
        m.AddBoundaryCondition(KINDS[int(k)], [int(node)], int(c), float(v))
    return m


def run_oracle_problem(prob, steps, threads=8):
    om = oracle_model_from_problem(prob)
    # the device computes dt = 0.5 * critical dt; the oracle does the same
    dt = prob.dt_factor * om.critical_dt()
    run = O.ExplicitRun(om, steps * dt, steps, 0, threads=threads, reaction_nodes=prob.top, reaction_coord=2)
    run.advance(steps)
    return run, dt


def run_gpu_problem(prob, steps, precision):
    m = prob.make_model(precision)
    integ = prob.make_integrator(m, num_steps=steps)
    integ.Integrate()
    return m, integ


@pytest.mark.parametrize("name", ["single_elem_native_neohookean", "simple_deformation_modes", "wave_in_bar",
                                  "brick_with_fibers", "notched_plate_native_neohookean"])
def test_reference_decks_fp64(hip_device, name):
    g = load_golden(name)
    nls, of = int(g["num_load_steps"]), int(g["output_frequency"])
    om = golden_oracle_model(g)
    orun = O.ExplicitRun(om, float(g["t_final"]), nls, of, float(g["t_initial"]), keep_ip=True, threads=8)
    m = gpu_model_from_golden(g)
    integ = N.ExplicitTimeIntegrator(m, float(g["t_final"]), nls, of, float(g["t_initial"])).Initialize()
    assert np.array_equal(m.GetScalarNodeData("lumped_mass"), orun.m)
    outs = output_steps(nls, of)
    labels = ("displacement", "velocity", "internal_force", "acceleration")
    refs = {k: [] for k in labels + ("F", "stress")}
    gots = {k: [] for k in labels + ("F", "stress")}

    def record(step, t):
        orun.advance(step + 1 - orun.step)
        assert t == orun.time
        for label, ref in zip(labels, (orun.u, orun.v, orun.f, orun.a)):
            refs[label].append(ref.copy())
            gots[label].append(m.GetVectorNodeData(label))
        refs["F"].append(orun.Fip.copy())
        gots["F"].append(m.GetIntegrationPointData("deformation_gradient"))
        refs["stress"].append(orun.sip.copy())
        gots["stress"].append(m.GetIntegrationPointData("stress"))

    integ.Integrate(callback=record)
    assert len(refs["displacement"]) == len(outs)
    # error at each output time relative to the field's largest magnitude over
    # the run (one absolute tolerance per variable, as the reference's exodiff
    # files set); early outputs carry strains of 1e-12..1e-8 whose stresses are
    # differences of O(1) numbers
    for k in refs:
        scale = max(np.abs(r).max() for r in refs[k])
        err = max(np.abs(g - r).max() for g, r in zip(gots[k], refs[k])) / scale
        print(f"{name} {k}: {err:.2e}")
        # the FP64 path keeps the reference's operation order and libm cube-root
        # method: every field is bit-identical to the oracle
        assert err == 0.0, (name, k, err)


@pytest.mark.parametrize("name", ["single_elem_native_neohookean", "simple_deformation_modes", "wave_in_bar",
                                  "brick_with_fibers", "notched_plate_native_neohookean"])
def test_reference_decks_mixed(hip_device, name):
    """The mixed path on the reference's own regression decks (two blocks in
    brick_with_fibers, an unstructured mesh in notched_plate, whose tiles sum
    by colour) against the FP64 oracle at every output time: u and F within
    1e-6, v, a, f_int and sigma within 1e-5 of each field's largest magnitude
    over the run (measured: u <= 1.6e-7, F <= 6.1e-8, v, a, f, sigma <=
    3.5e-6, largest on wave_in_bar over its 1000 steps)."""
    g = load_golden(name)
    nls, of = int(g["num_load_steps"]), int(g["output_frequency"])
    om = golden_oracle_model(g)
    orun = O.ExplicitRun(om, float(g["t_final"]), nls, of, float(g["t_initial"]), keep_ip=True, threads=8)
    m = gpu_model_from_golden(g, "mixed")
    integ = N.ExplicitTimeIntegrator(m, float(g["t_final"]), nls, of, float(g["t_initial"])).Initialize()
    labels = ("displacement", "velocity", "internal_force", "acceleration")
    refs = {k: [] for k in labels + ("F", "stress")}
    gots = {k: [] for k in labels + ("F", "stress")}

    def record(step, t):
        orun.advance(step + 1 - orun.step)
        for label, ref in zip(labels, (orun.u, orun.v, orun.f, orun.a)):
            refs[label].append(ref.copy())
            gots[label].append(m.GetVectorNodeData(label))
        refs["F"].append(orun.Fip.copy())
        gots["F"].append(m.GetIntegrationPointData("deformation_gradient"))
        refs["stress"].append(orun.sip.copy())
        gots["stress"].append(m.GetIntegrationPointData("stress"))

    integ.Integrate(callback=record)
    m.close()
    bounds = {"displacement": 1e-6, "F": 1e-6, "velocity": 1e-5, "acceleration": 1e-5, "internal_force": 1e-5,
              "stress": 1e-5}
    for k, bound in bounds.items():
        scale = max(np.abs(r).max() for r in refs[k])
        err = max(np.abs(g_ - r).max() for g_, r in zip(gots[k], refs[k])) / scale
        print(f"{name} {k}: {err:.2e}")
        assert err <= bound, (name, k, err)


@pytest.mark.parametrize("ref_table", ["1", "0"])
def test_cube16_fp64_1000_steps(hip_device, monkeypatch, ref_table):
    """1000 FP64 steps bit-identical to the oracle, with the reference inverse
    Jacobians read from the table formed at setup and formed in every launch
    (NSM_FP64_REF_TABLE; small meshes use the table by default)."""
    monkeypatch.setenv("NSM_FP64_REF_TABLE", ref_table)
    prob = N.workload("cube16_fp64")
    steps = prob.num_steps
    m, integ = run_gpu_problem(prob, steps, "fp64")
    run, dt = run_oracle_problem(prob, steps)
    assert integ.dt_user == run.dt_user
    assert integ.time_current == run.time
    for label, ref in (("displacement", run.u), ("velocity", run.v), ("internal_force", run.f)):
        got = m.GetVectorNodeData(label)
        assert np.array_equal(got, ref), (label, rel_err(got, ref))  # bit-identical (target: 1e-12)
    R = m.GetReactionHistory(0, steps)
    # the reaction is a parallel tree sum on the device (a serial sum in the oracle)
    assert max_rel_err(R, run.reaction) <= 1e-14
    assert np.abs(R).max() > 0


@pytest.mark.parametrize("n,ref_table", [(8, "1"), (9, "1"), (9, "0")])
def test_one_step_internal_force_two_blocks(hip_device, monkeypatch, n, ref_table):
    """Forces, F and sigma of one evaluation bit-identical to the oracle; n = 9
    (729 elements) ends the FP64 element groups of 8 with a partial one and
    starts the second block inside a group."""
    monkeypatch.setenv("NSM_FP64_REF_TABLE", ref_table)
    prob = N.CubeProblem(n=n, materials=[N.problems.STEEL, N.problems.ALUMINIUM], num_steps=1, seed=77)
    m = prob.make_model("fp64")
    rng = np.random.default_rng(2)
    u = rng.normal(0, 1e-4, (prob.num_nodes, 3))
    m.SetVectorNodeData("displacement", u)
    m.ComputeInternalForce()
    om = oracle_model_from_problem(prob)
    f, Fip, sip = om.internal_force(u, want_ip=True)
    assert np.array_equal(m.GetVectorNodeData("internal_force"), f)
    assert np.array_equal(m.GetIntegrationPointData("deformation_gradient"), Fip)
    assert np.array_equal(m.GetIntegrationPointData("stress"), sip)


def test_small_and_empty_blocks(hip_device):
    # one element cube, plus a model with an empty middle block
    for prob in (N.CubeProblem(n=1, materials=[N.problems.STEEL], num_steps=20),
                 N.CubeProblem(n=3, materials=[N.problems.STEEL, N.problems.TITANIUM, N.problems.POLYMER],
                               num_steps=20)):
        m, integ = run_gpu_problem(prob, prob.num_steps, "fp64")
        run, _ = run_oracle_problem(prob, prob.num_steps, threads=1)
        assert np.array_equal(m.GetVectorNodeData("displacement"), run.u)
    X, conn = N.jittered_cube(2, 3, 0.1)
    blocks = [N.Block(1, conn[:4], N.NeohookeanMaterial(7800, 1e11, 5e10)),
              N.Block(2, conn[:0], N.NeohookeanMaterial(2700, 7e10, 3e10)),
              N.Block(3, conn[4:], N.NeohookeanMaterial(7800, 1e11, 5e10))]
    m = N.ModelData(X, blocks)
    m.ComputeLumpedMass()
    om = O.Model(X, conn, [0, 4, 4, 8], [7800, 2700, 7800], [1e11, 7e10, 1e11], [5e10, 3e10, 5e10])
    assert np.array_equal(m.GetScalarNodeData("lumped_mass"), om.lumped_mass())
    assert m.GetCriticalTimeStep() == om.critical_dt()


def test_fp64_runs_are_deterministic(hip_device):
    prob = N.CubeProblem(n=24, materials=[N.problems.STEEL, N.problems.ALUMINIUM], num_steps=50)
    outs = []
    for _ in range(2):
        m, _ = run_gpu_problem(prob, prob.num_steps, "fp64")
        outs.append((m.GetVectorNodeData("displacement"), m.GetVectorNodeData("internal_force"),
                     m.GetReactionHistory(0, prob.num_steps)))
        m.close()
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


def test_mixed_precision_cube16_against_fp64_oracle(hip_device):
    """Stated mixed-precision error on the 16^3 workload after 1000 steps:
    displacement <= 1e-6, internal force <= 1e-5, reaction history <= 1e-5
    (norm-wise relative to the FP64 oracle)."""
    prob = N.workload("cube16_fp64")
    steps = prob.num_steps
    m, integ = run_gpu_problem(prob, steps, "mixed")
    run, _ = run_oracle_problem(prob, steps)
    eu = rel_err(m.GetVectorNodeData("displacement"), run.u)
    ef = rel_err(m.GetVectorNodeData("internal_force"), run.f)
    ev = rel_err(m.GetVectorNodeData("velocity"), run.v)
    eR = rel_err(m.GetReactionHistory(0, steps), run.reaction)
    print(f"mixed vs fp64 oracle: u {eu:.2e} v {ev:.2e} f {ef:.2e} R {eR:.2e}")
    assert eu <= 1e-6 and ef <= 1e-5 and eR <= 1e-5 and ev <= 1e-5


def test_mixed_runs_are_deterministic(hip_device):
    prob = N.CubeProblem(n=20, materials=[N.problems.STEEL], num_steps=30)
    a, _ = run_gpu_problem(prob, prob.num_steps, "mixed")
    b, _ = run_gpu_problem(prob, prob.num_steps, "mixed")
    assert np.array_equal(a.GetVectorNodeData("internal_force"), b.GetVectorNodeData("internal_force"))
    assert np.array_equal(a.GetReactionHistory(0, 30), b.GetReactionHistory(0, 30))


def test_large_cube_properties(hip_device):
    """128^3, two blocks (the benchmark mesh), its full 500 steps:
    size-independent checks, and mixed precision against the FP64 run on the
    device: u <= 1e-6, v, f_int and the reaction history <= 1e-5 (norm-wise
    relative); u and v also over the nodes without prescribed conditions."""
    prob = N.workload("cube128_2blk")
    steps = prob.num_steps
    res = {}
    for prec in ("fp64", "mixed"):
        print(f"cube128_2blk {prec}: {steps} steps", flush=True)
        m, integ = run_gpu_problem(prob, steps, prec)
        f = m.GetVectorNodeData("internal_force")
        u = m.GetVectorNodeData("displacement")
        v = m.GetVectorNodeData("velocity")
        R = m.GetReactionHistory(0, steps)
        # translation invariance: internal forces sum to zero
        assert np.abs(f.sum(axis=0)).max() <= 1e-9 * np.abs(f).sum(axis=0).max()
        # prescribed top velocity: u_z on the top face = v_top * t exactly as accumulated
        assert np.allclose(u[prob.top, 2], integ.time_current * prob.v_top, rtol=1e-12)
        assert np.all(u[prob.bottom] == 0.0)
        res[prec] = (u, v, f, R)
        m.close()
        del m
    free = np.ones(prob.num_nodes, bool)
    free[prob.top] = free[prob.bottom] = False
    errs = {name: rel_err(res["mixed"][i], res["fp64"][i]) for i, name in enumerate(("u", "v", "f", "R"))}
    errs["u_free"] = rel_err(res["mixed"][0][free], res["fp64"][0][free])
    errs["v_free"] = rel_err(res["mixed"][1][free], res["fp64"][1][free])
    print("cube128_2blk mixed vs fp64 GPU after %d steps: " % steps +
          " ".join(f"{k} {e:.2e}" for k, e in errs.items()), flush=True)
    assert errs["u"] <= 1e-6 and errs["u_free"] <= 1e-6
    assert errs["v"] <= 1e-5 and errs["v_free"] <= 1e-5
    assert errs["f"] <= 1e-5 and errs["R"] <= 1e-5


@pytest.mark.parametrize("case", ["uniaxial", "simple shear", "general"])
def test_reaction_closed_form_on_device(hip_device, case):
    """One device step from a uniform affine state x = F X of an undistorted
    cube at rest: the top-face reaction equals the traction resultant
    e_z . sigma(F) . J F^{-T} e_Z (the oracle's stress) to 1e-12 of its
    magnitude in FP64 and 1e-5 in mixed precision, and the bottom face's is
    its negative (one material; the same closed form pins the oracle)."""
    from tests.test_oracle_math import AFFINE, STEEL_KG, closed_form_reaction, traction_resultant

    F = AFFINE[case]
    n = 8
    X, conn = N.jittered_cube(n, 1, 0.0)
    u = X @ (F - np.eye(3)).T
    n1 = n + 1
    top = (np.arange(n1 * n1) + n * n1 * n1).astype(np.int32)
    bottom = np.arange(n1 * n1, dtype=np.int32)
    scale = np.linalg.norm(traction_resultant(F, STEEL_KG[1], STEEL_KG[2]))
    for precision, tol in (("fp64", 1e-12), ("mixed", 1e-5)):
        for nodes, axis, sign in ((top, 2, 1.0), (bottom, 2, -1.0), (top, 0, 1.0)):
            m = N.ModelData(X, [N.Block(1, conn, N.NeohookeanMaterial(*STEEL_KG))], precision)
            m.SetReactionNodes(nodes, axis)
            m.SetVectorNodeData("displacement", u)
            m.ComputeLumpedMass()
            m.ExplicitSteps(0, 1, 1, 0, 1e-9, 0.0)
            R = m.GetReactionHistory(0, 1)[0]
            ref = sign * closed_form_reaction(F, STEEL_KG[1], STEEL_KG[2], axis)
            assert abs(R - ref) <= tol * scale, (precision, case, axis, R, ref)
            m.close()


def test_velocity_then_displacement_condition_on_one_node(hip_device):
    """A prescribed displacement listed after a prescribed velocity on the
    same node and coordinate: the reference applies both in order
    (src/nimble_boundary_condition_manager.h:136-217), so at delta_t = 0 the
    velocity condition holds, and UQ samples (which take velocities only)
    keep it for the whole run.  FP64 device = oracle, bit for bit."""
    prob = N.CubeProblem(n=4, materials=[N.problems.STEEL], num_steps=12, v_top=0.0)
    prob.build()
    side = np.arange(0, prob.num_nodes, 7, dtype=np.int32)  # some nodes of every layer
    bcs = prob.bc_list() + [("prescribed_velocity", side, 0, 0.75), ("prescribed_displacement", side, 0, 2e-6),
                            ("prescribed_velocity", side, 1, -0.5)]
    kinds = {"initial_velocity": 0, "prescribed_velocity": 1, "prescribed_displacement": 2}

    def gpu_model():
        m = N.ModelData(prob.X, prob.blocks(), "fp64")
        for kind, nodes, c, mag in bcs:
            m.AddBoundaryCondition(kind, nodes, c, mag)
        m.SetReactionNodes(prob.top, 2)
        return m

    om = oracle_model_from_problem(prob)
    bn, bc, bk, bv = [], [], [], []
    for kind, nodes, c, mag in bcs:
        for nd in nodes:
            bn.append(nd), bc.append(c), bk.append(kinds[kind]), bv.append(mag)
    om = O.Model(om.X, om.conn, om.blk_off, om.rho, om.K, om.G, np.array(bn, np.int32), np.array(bc, np.int32),
                 np.array(bk, np.int32), np.array(bv))
    m = gpu_model()
    integ = N.ExplicitTimeIntegrator(m, 12 * 1e-7, 12, 5).Initialize()
    v0 = m.GetVectorNodeData("velocity")
    assert np.all(v0[side, 0] == 0.75) and np.all(v0[side, 1] == -0.5)
    run = O.ExplicitRun(om, 12 * 1e-7, 12, 5)
    assert np.array_equal(v0, run.v)
    integ.Integrate()
    run.advance(12)
    for label, ref in (("displacement", run.u), ("velocity", run.v), ("internal_force", run.f)):
        assert np.array_equal(m.GetVectorNodeData(label), ref), label
    m.close()
    # UQ: one exact sample with the nominal moduli and one approximate sample
    m = gpu_model()
    K, G = prob.blocks()[0].material.bulk_modulus, prob.blocks()[0].material.shear_modulus
    m.SetUqSamples([[K]], [[G]], [[0.0, 1.0]])
    integ = N.ExplicitTimeIntegrator(m, 12 * 1e-7, 12, 5).Initialize()
    integ.Integrate()
    uq = O.UqRun(om, [[K]], [[G]], [[0.0, 1.0]], 12 * 1e-7, 12, 5)
    uq.advance(12)
    for s in range(2):
        ru, rv, rf = uq.sample(s)
        assert np.all(m.GetUqNodeData(s, "velocity")[side, 0] == 0.75)
        for label, ref in (("displacement", ru), ("velocity", rv), ("internal_force", rf)):
            assert np.array_equal(m.GetUqNodeData(s, label), ref), (s, label)
    m.close()


def test_largest_cube_mixed_against_fp64_gpu(hip_device):
    """256^3, four z-layered blocks, the full 200 steps: mixed precision
    against an FP64 run on the device, with the size-independent checks."""
    prob = N.workload("cube256_4blk")
    steps = prob.num_steps
    res = {}
    for prec in ("fp64", "mixed"):
        print(f"cube256_4blk {prec}: {steps} steps", flush=True)
        m, integ = run_gpu_problem(prob, steps, prec)
        f = m.GetVectorNodeData("internal_force")
        u = m.GetVectorNodeData("displacement")
        R = m.GetReactionHistory(0, steps)
        assert np.abs(f.sum(axis=0)).max() <= 1e-9 * np.abs(f).sum(axis=0).max()
        assert np.allclose(u[prob.top, 2], integ.time_current * prob.v_top, rtol=1e-12)
        assert np.all(u[prob.bottom] == 0.0)
        res[prec] = (u, f, R)
        m.close()
        del m
    eu, ef, eR = (rel_err(res["mixed"][i], res["fp64"][i]) for i in range(3))
    print(f"cube256_4blk mixed vs fp64 GPU after {steps} steps: u {eu:.2e} f {ef:.2e} R {eR:.2e}")
    assert eu <= 1e-6 and ef <= 1e-5 and eR <= 1e-5


def test_reaction_over_scattered_nodes(hip_device):
    """Reaction sum over an arbitrary node set (many node tiles, some nodes
    with zero force in early steps) against the oracle's serial sum."""
    prob = N.CubeProblem(n=12, materials=[N.problems.STEEL, N.problems.ALUMINIUM], num_steps=40)
    rng = np.random.default_rng(3)
    nodes = np.sort(rng.choice(prob.num_nodes, 300, replace=False)).astype(np.int32)
    for precision, tol in (("fp64", 1e-13), ("mixed", 1e-4)):
        m = prob.make_model(precision)
        m.SetReactionNodes(nodes, 0)
        integ = prob.make_integrator(m)
        integ.Integrate()
        om = oracle_model_from_problem(prob)
        dt = prob.dt_factor * om.critical_dt()
        run = O.ExplicitRun(om, prob.num_steps * dt, prob.num_steps, 0, reaction_nodes=nodes, reaction_coord=0)
        run.advance(prob.num_steps)
        R = m.GetReactionHistory(0, prob.num_steps)
        scale = np.abs(run.reaction).max()
        assert np.abs(R - run.reaction).max() <= tol * scale, (precision, np.abs(R - run.reaction).max() / scale)
        m.close()


def test_sampled_reaction_history(hip_device):
    """64-point QoI on the device: identical to the oracle's sampling rule
    applied to the device history; FP64 history sampled = oracle's sampled
    history to 1e-13, mixed to 1e-5."""
    prob = N.CubeProblem(n=10, materials=[N.problems.STEEL, N.problems.ALUMINIUM], num_steps=100)
    run, _ = run_oracle_problem(prob, prob.num_steps)
    ref = O.sample_history(run.reaction, 64)
    for precision, tol in (("fp64", 1e-13), ("mixed", 1e-5)):
        m, integ = run_gpu_problem(prob, prob.num_steps, precision)
        R = m.GetReactionHistory(0, prob.num_steps)
        s = m.SampleReactionHistory(prob.num_steps, 64)
        assert np.array_equal(s, O.sample_history(R, 64))
        assert max_rel_err(s, ref) <= tol, (precision, max_rel_err(s, ref))
        m.close()


@pytest.mark.parametrize("mode", ["fp64", "fp64-no-table", "tiles", "slots"])
def test_inverted_element_in_model_reports_error_and_recovers(hip_device, monkeypatch, mode):
    """A displacement that turns elements inside out makes the model's force
    evaluation fail with NSM_ERR_NEG_JACOBIAN (the reference's Invert3x3
    check); the status is cleared, so a valid displacement then works."""
    if mode == "fp64-no-table":
        monkeypatch.setenv("NSM_FP64_REF_TABLE", "0")
        mode = "fp64"
    if mode != "fp64":
        monkeypatch.setenv("NSM_MIXED_ASSEMBLY", mode)
    prob = N.CubeProblem(n=4, materials=[N.problems.STEEL], num_steps=1)
    m = prob.make_model("fp64" if mode == "fp64" else "mixed")
    assert m.AssemblyInfo()["tiled"] == (mode == "tiles")
    u = np.zeros((prob.num_nodes, 3))
    center = 2 + 5 * (2 + 5 * 2)  # interior node (2, 2, 2)
    u[center, 2] = -0.6  # pushed through the node below it (h = 0.25)
    m.SetVectorNodeData("displacement", u)
    with pytest.raises(N.NimbleError) as ei:
        m.ComputeInternalForce()
    assert ei.value.status == 3
    m.SetVectorNodeData("displacement", np.zeros_like(u))
    m.ComputeInternalForce()
    assert np.abs(m.GetVectorNodeData("internal_force")).max() < 1e-3  # undeformed: f ~ 0
    m.close()
