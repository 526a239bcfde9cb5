"""Batched ensembles on the device.

Every sample of an ensemble must follow the trajectory a ModelData holding
that sample alone computes: bit for bit in both precisions (same element
arithmetic, same summation orders, same reaction sums), including the lumped
mass and the critical time step, which the ensemble computes per sample.  The
single-model FP64 path is itself bit-identical to the oracle (test_gpu_model);
one sample is also checked against the oracle directly here.  Besides
jittered cubes, the meshes of three reference decks (one unstructured, one
with two blocks) run side by side in one ensemble.  Graph replays
must equal eager runs, also after the samples' parameters change.  The full
ensemble workload (256 samples on the 32^3 cube and 64 on uneven cubes, 400
steps) is checked in mixed precision against the same ensemble in FP64.
"""
import os

import numpy as np
import pytest

import nimblesm_amd as N
from nimblesm_amd.ensemble import EnsembleProblem, config2_draws, draw_parameters
from oracle import oracle as O
from tests.helpers import oracle_model_from_problem, rel_err

pytestmark = pytest.mark.gpu

FIELDS = ("displacement", "velocity", "acceleration", "internal_force")


def single_model(n, materials, v_top, steps, precision, tile_elems):
    """A CubeProblem run of one sample alone (the plain one-at-a-time solve)."""
    prob = N.CubeProblem(n=n, materials=materials, num_steps=steps, v_top=v_top, name="single")
    old = os.environ.get("NSM_TILE_ELEMS")
    os.environ["NSM_TILE_ELEMS"] = str(tile_elems)
    try:
        m = prob.make_model(precision)
    finally:
        if old is None:
            del os.environ["NSM_TILE_ELEMS"]
        else:
            os.environ["NSM_TILE_ELEMS"] = old
    integ = prob.make_integrator(m)
    integ.Integrate()
    return prob, m


def two_block_ensemble(precision, tile_elems, steps):
    """Meshes 5^3 (one block) and 6^3 (two z-layered blocks), three samples
    each with their own materials and top velocity."""
    ens = N.Ensemble(precision, tile_elems)
    rng = np.random.default_rng(7)
    cases = []
    for n, nb in ((5, 1), (6, 2)):
        prob = N.CubeProblem(n=n, materials=[N.problems.STEEL] * nb, num_steps=steps).build()
        mid = ens.AddMesh(prob.X, [b.conn for b in prob.blocks()])
        for kind, nodes, c, _ in prob.bc_list():
            ens.AddBoundaryCondition(mid, kind, nodes, c)
        ens.SetReactionNodes(mid, prob.top, 2)
        mats, vs = [], []
        for _ in range(3):
            ms = [(E, nu, rho) for E, nu, rho in zip(rng.uniform(50e9, 250e9, nb), rng.uniform(0.22, 0.42, nb),
                                                     rng.uniform(2000, 9000, nb))]
            mats.append(ms)
            vs.append(float(rng.uniform(0.5, 2.0)))
        nm = [[N.NeohookeanMaterial.from_young(rho, E, nu) for E, nu, rho in ms] for ms in mats]
        mag = np.zeros((3, 4))
        mag[:, 3] = vs
        ens.AddSamples(mid, [[x.density for x in r] for r in nm], [[x.bulk_modulus for x in r] for r in nm],
                       [[x.shear_modulus for x in r] for r in nm], mag)
        cases += [(n, ms, v) for ms, v in zip(mats, vs)]
    ens.Setup()
    return ens, cases


@pytest.mark.parametrize("precision,tile", [("fp64", 0), ("mixed", 64), ("mixed", 128)])
def test_samples_match_single_models(hip_device, precision, tile):
    steps = 40
    ens, cases = two_block_ensemble(precision, tile, steps)
    dts = ens.CriticalTimeSteps()
    ft = steps * (0.5 * dts)
    ens.Run(steps, ft, use_graph=False)
    for s, (n, mats, v) in enumerate(cases):
        prob, m = single_model(n, mats, v, steps, precision, tile or 512)
        assert dts[s] == m.GetCriticalTimeStep(), s
        assert np.array_equal(ens.GetNodeData(s, "lumped_mass"), m.GetScalarNodeData("lumped_mass")), s
        for f in FIELDS:
            assert np.array_equal(ens.GetNodeData(s, f), m.GetVectorNodeData(f)), (s, f)
        assert np.array_equal(ens.GetReactionHistory(s), m.GetReactionHistory(0, steps)), s
        q = ens.SampleReactionHistory(64)[s]
        assert np.array_equal(q, m.SampleReactionHistory(steps, 64)), s
        m.close()
    ens.close()


def test_fp64_sample_matches_oracle(hip_device):
    steps = 60
    ens, cases = two_block_ensemble("fp64", 0, steps)
    dts = ens.CriticalTimeSteps()
    ens.Run(steps, steps * (0.5 * dts), use_graph=True)
    for s in (1, 4):
        n, mats, v = cases[s]
        prob = N.CubeProblem(n=n, materials=mats, num_steps=steps, v_top=v).build()
        om = oracle_model_from_problem(prob)
        assert dts[s] == om.critical_dt()
        run = O.ExplicitRun(om, steps * (0.5 * om.critical_dt()), steps, 0, reaction_nodes=prob.top,
                            reaction_coord=2)
        run.advance(steps)
        assert np.array_equal(ens.GetNodeData(s, "displacement"), run.u)
        assert np.array_equal(ens.GetNodeData(s, "velocity"), run.v)
        assert np.array_equal(ens.GetNodeData(s, "internal_force"), run.f)
        assert rel_err(ens.GetReactionHistory(s), run.reaction) <= 1e-14
    ens.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_replays_equal_eager_runs(hip_device, precision):
    import torch

    steps = 25
    d = draw_parameters(6, seed=3, n=(4, 6))
    prob = EnsembleProblem(d, num_steps=steps, precision=precision, tile_elems=64).build()
    ens = prob.ens
    eager = prob.run(use_graph=False)
    u_eager = ens.GetNodeData(2, "displacement")
    for k in range(3):
        # unrelated eager device work and synchronizations between replays
        x = torch.ones(1 << 20, device="cuda") * (k + 1)
        torch.cuda.synchronize()
        assert float(x.sum()) == float((k + 1) << 20)
        got = prob.run(use_graph=True)
        assert np.array_equal(got, eager), k
        assert np.array_equal(ens.GetNodeData(2, "displacement"), u_eager), k
    # new parameters: a replay of the captured run uses them
    d2 = draw_parameters(6, seed=4, n=(4, 6))
    for mid, n in enumerate(sorted(set(d.n))):
        idx = np.nonzero(prob.draws.n == n)[0]
        new = d2.subset(np.arange(len(idx)))
        mats = [N.NeohookeanMaterial.from_young(new.rho[i], new.E[i], new.nu[i]) for i in range(len(idx))]
        mag = np.zeros((len(idx), 4))
        mag[:, 3] = new.v_top
        first = int(np.nonzero(np.array(ens.sample_mesh) == mid)[0][0])
        ens.SetSamples(first, [m.density for m in mats], [m.bulk_modulus for m in mats],
                       [m.shear_modulus for m in mats], mag)
    eager2 = prob.run(use_graph=False)
    assert not np.array_equal(eager2, eager)
    for k in range(2):
        assert np.array_equal(prob.run(use_graph=True), eager2), k
    prob.close()


def test_reduced_uneven_ensemble_mixed_against_fp64(hip_device):
    d = draw_parameters(12, seed=11, n=(4, 6, 8, 10))
    qs = {}
    for precision in ("fp64", "mixed"):
        prob = EnsembleProblem(d, num_steps=120, precision=precision, tile_elems=64).build()
        qs[precision] = prob.run()
        prob.close()
    ref = qs["fp64"]
    assert np.isfinite(ref).all() and (np.abs(ref).max(axis=1) > 0).all()
    per_sample = np.linalg.norm(qs["mixed"] - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert per_sample.max() <= 1e-5, per_sample
    assert rel_err(qs["mixed"].mean(axis=0), ref.mean(axis=0)) <= 1e-5


def test_full_ensemble_workload(hip_device):
    """256 samples on 32^3 + 64 on uneven cubes, 400 steps: mixed against
    FP64 (mean QoI and every sample), and samples of every edge count
    bit-identical to one-at-a-time solves of the plain path."""
    d = config2_draws()
    res = {}
    for precision in ("fp64", "mixed"):
        prob = EnsembleProblem(d, num_steps=400, precision=precision, tile_elems=512).build()
        res[precision] = (prob.run(), prob.ens.CriticalTimeSteps()[np.argsort(prob.order)])
        prob.close()
    ref, q = res["fp64"][0], res["mixed"][0]
    mean_err = rel_err(q.mean(axis=0), ref.mean(axis=0))
    per_sample = np.linalg.norm(q - ref, axis=1) / np.linalg.norm(ref, axis=1)
    print(f"mixed vs FP64: mean-QoI error {mean_err:.3e}, worst sample {per_sample.max():.3e}", flush=True)
    assert mean_err <= 1e-3 and per_sample.max() <= 1e-5
    # one sample of each edge count against the plain path on its own
    for n in sorted(set(d.n)):
        i = int(np.nonzero(d.n == n)[0][0])
        for precision in ("fp64", "mixed"):
            prob1, m = single_model(n, [(d.E[i], d.nu[i], d.rho[i])], d.v_top[i], 400, precision, 512)
            assert m.GetCriticalTimeStep() == res[precision][1][i]
            assert np.array_equal(m.SampleReactionHistory(400, 64), res[precision][0][i]), (n, precision)
            m.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_displacement_and_initial_velocity_conditions(hip_device, precision):
    """Per-sample magnitudes of an initial velocity, a prescribed
    displacement, and a prescribed velocity followed by a prescribed
    displacement on the same nodes, with output steps every 7 steps: each
    sample bit-identical to a ModelData run with those conditions."""
    n, steps, of = 6, 30, 7
    X, conn = N.jittered_cube(n, 99, 0.1)
    n1 = n + 1
    ids = np.arange(n1 ** 3, dtype=np.int32)
    k = ids // (n1 * n1)
    j = (ids // n1) % n1
    bottom, top, side = ids[k == 0], ids[k == n], ids[(j == 0) & (k > 0) & (k < n)]
    conds = [("initial_velocity", ids, 0), ("prescribed_velocity", bottom, 0), ("prescribed_velocity", bottom, 1),
             ("prescribed_velocity", bottom, 2), ("prescribed_displacement", top, 2),
             ("prescribed_velocity", side, 1), ("prescribed_displacement", side, 1)]
    rng = np.random.default_rng(12)
    S = 3
    mats = [N.NeohookeanMaterial.from_young(rng.uniform(5000, 9000), rng.uniform(100e9, 250e9),
                                            rng.uniform(0.25, 0.4)) for _ in range(S)]
    mags = np.zeros((S, len(conds)))
    mags[:, 0] = rng.uniform(-0.5, 0.5, S)
    mags[:, 4] = rng.uniform(1e-5, 3e-5, S)
    mags[:, 5] = rng.uniform(0.1, 0.3, S)
    mags[:, 6] = rng.uniform(-2e-5, 2e-5, S)
    ens = N.Ensemble(precision, 64)
    mid = ens.AddMesh(X, [conn])
    for kind, nodes, c in conds:
        ens.AddBoundaryCondition(mid, kind, nodes, c)
    ens.SetReactionNodes(mid, top, 2)
    ens.AddSamples(mid, [m.density for m in mats], [m.bulk_modulus for m in mats], [m.shear_modulus for m in mats],
                   mags)
    ens.Setup()
    dts = ens.CriticalTimeSteps()
    ens.Run(steps, steps * (0.5 * dts), output_frequency=of, use_graph=True)
    os.environ["NSM_TILE_ELEMS"] = "64"
    try:
        for s in range(S):
            m = N.ModelData(X, [N.Block(1, conn, mats[s])], precision)
            for (kind, nodes, c), v in zip(conds, mags[s]):
                m.AddBoundaryCondition(kind, nodes, c, float(v))
            m.SetReactionNodes(top, 2)
            m.ComputeLumpedMass()
            assert m.GetCriticalTimeStep() == dts[s]
            integ = N.ExplicitTimeIntegrator(m, steps * (0.5 * dts[s]), steps, of).Initialize()
            integ.Integrate()
            for f in FIELDS:
                assert np.array_equal(ens.GetNodeData(s, f), m.GetVectorNodeData(f)), (s, f)
            assert np.array_equal(ens.GetReactionHistory(s), m.GetReactionHistory(0, steps)), s
            m.close()
    finally:
        del os.environ["NSM_TILE_ELEMS"]
    ens.close()


def deck_conditions(g):
    """The golden deck's conditions as (kind, nodes, coordinate, value)
    groups of consecutive entries with one kind, coordinate and value."""
    kinds = ("initial_velocity", "prescribed_velocity", "prescribed_displacement")
    key = list(zip(g["bc_kind"], g["bc_coord"], g["bc_value"]))
    out, i = [], 0
    while i < len(key):
        j = i
        while j < len(key) and key[j] == key[i]:
            j += 1
        kind, c, v = key[i]
        out.append((kinds[int(kind)], g["bc_node"][i:j].astype(np.int32), int(c), float(v)))
        i = j
    return out


@pytest.mark.parametrize("precision,tile", [("fp64", 0), ("mixed", 64), ("mixed", 128)])
def test_reference_deck_meshes_in_one_ensemble(hip_device, precision, tile):
    """The meshes of three reference decks (wave_in_bar; brick_with_fibers,
    two blocks; notched_plate, unstructured) in one ensemble, two samples
    each: the deck's own parameters, and moduli 0.8x, density 1.1x with
    condition magnitudes 0.7x.  Every sample is bit-identical to a ModelData
    run of the same mesh, materials and conditions (deck time step, 40 steps,
    output steps every 6)."""
    from tests.conftest import load_golden
    steps, of = 40, 6
    names = ("wave_in_bar", "brick_with_fibers", "notched_plate_native_neohookean")
    ens = N.Ensemble(precision, tile)
    cases = []
    for name in names:
        g = load_golden(name)
        blocks = [g["conn"][int(g["blk_off"][b]):int(g["blk_off"][b + 1])] for b in range(len(g["rho"]))]
        conds = deck_conditions(g)
        react = next(c for c in conds if c[0] == "prescribed_velocity")  # the held end, x reaction
        mid = ens.AddMesh(g["X"], blocks)
        for kind, nodes, c, _ in conds:
            ens.AddBoundaryCondition(mid, kind, nodes, c)
        ens.SetReactionNodes(mid, react[1], react[2])
        rho, K, G = (np.asarray(g[k], np.float64) for k in ("rho", "K", "G"))
        mags = np.array([v for *_, v in conds])
        params = [(rho, K, G, mags), (1.1 * rho, 0.8 * K, 0.8 * G, 0.7 * mags)]
        ens.AddSamples(mid, [p[0] for p in params], [p[1] for p in params], [p[2] for p in params],
                       [p[3] for p in params])
        dt = float(g["t_final"]) / int(g["num_load_steps"])
        cases += [(g, blocks, conds, react, p, dt) for p in params]
    ens.Setup()
    dts = ens.CriticalTimeSteps()
    ens.Run(steps, np.array([c[-1] * steps for c in cases]), output_frequency=of, use_graph=True)
    old = os.environ.get("NSM_TILE_ELEMS")
    os.environ["NSM_TILE_ELEMS"] = str(tile or 512)
    try:
        for s, (g, blocks, conds, react, (rho, K, G, mags), dt) in enumerate(cases):
            m = N.ModelData(g["X"], [N.Block(int(g["blk_ids"][b]), blocks[b],
                                             N.NeohookeanMaterial(rho[b], K[b], G[b])) for b in range(len(blocks))],
                            precision)
            for (kind, nodes, c, _), v in zip(conds, mags):
                m.AddBoundaryCondition(kind, nodes, c, float(v))
            m.SetReactionNodes(react[1], react[2])
            integ = N.ExplicitTimeIntegrator(m, dt * steps, steps, of).Initialize()
            assert integ.critical_time_step == dts[s], s
            integ.Integrate()
            assert np.array_equal(ens.GetNodeData(s, "lumped_mass"), m.GetScalarNodeData("lumped_mass")), s
            for f in FIELDS:
                assert np.array_equal(ens.GetNodeData(s, f), m.GetVectorNodeData(f)), (s, f)
            assert np.array_equal(ens.GetReactionHistory(s), m.GetReactionHistory(0, steps)), s
            m.close()
    finally:
        if old is None:
            del os.environ["NSM_TILE_ELEMS"]
        else:
            os.environ["NSM_TILE_ELEMS"] = old
    ens.close()
