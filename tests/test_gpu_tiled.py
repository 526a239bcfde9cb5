"""Tile-local assembly of the mixed-precision path against the FP64 oracle
and against the element-slot + node-gather mixed path.

Stated bounds (norm-wise relative to the FP64 oracle): one evaluation of the
internal force <= 1e-5, F <= 1e-6, sigma <= 1e-5; explicit runs: displacement
<= 1e-6, velocity, internal force and reaction history <= 1e-5.  Between the
tiled and the slot mixed paths the same FP32 element forces are summed in
another order, and the tiled path stores the partial sums of nodes shared
between tiles in FP32: internal force within 1e-6 for one evaluation.
"""
import numpy as np
import pytest

import nimblesm_amd as N
from oracle import oracle as O
from tests.helpers import oracle_model_from_problem, rel_err

pytestmark = pytest.mark.gpu


def mesh_model(X, conn, blk_off, mats, precision, bcs, reaction=None):
    blocks = [N.Block(b + 1, conn[blk_off[b]:blk_off[b + 1]], mats[b]) for b in range(len(mats))]
    m = N.ModelData(X, blocks, precision)
    for kind, nodes, c, v in bcs:
        m.AddBoundaryCondition(kind, nodes, c, v)
    if reaction is not None:
        m.SetReactionNodes(reaction, 2)
    return m


def oracle_model(X, conn, blk_off, mats, bcs):
    kinds = {"initial_velocity": 0, "prescribed_velocity": 1, "prescribed_displacement": 2}
    bn, bc, bk, bv = [], [], [], []
    for kind, nodes, c, v in bcs:
        for n in nodes:
            bn.append(n), bc.append(c), bk.append(kinds[kind]), bv.append(v)
    return O.Model(X, conn, np.asarray(blk_off), np.array([m.density for m in mats]),
                   np.array([m.bulk_modulus for m in mats]), np.array([m.shear_modulus for m in mats]),
                   np.array(bn, np.int32), np.array(bc, np.int32), np.array(bk, np.int32), np.array(bv))


def cube_bcs(X, v_top=1.0):
    z = X[:, 2]
    bottom = np.nonzero(z == z.min())[0].astype(np.int32)
    top = np.nonzero(z == z.max())[0].astype(np.int32)
    bcs = [("prescribed_velocity", bottom, c, 0.0) for c in range(3)]
    bcs.append(("prescribed_velocity", top, 2, v_top))
    return bcs, top


def run_both(X, conn, blk_off, mats, steps, precision="mixed"):
    bcs, top = cube_bcs(X)
    m = mesh_model(X, conn, blk_off, mats, precision, bcs, top)
    om = oracle_model(X, conn, blk_off, mats, bcs)
    m.ComputeLumpedMass()
    dt = 0.5 * m.GetCriticalTimeStep()
    integ = N.ExplicitTimeIntegrator(m, steps * dt, steps, 0).Initialize()
    integ.Advance(steps)
    run = O.ExplicitRun(om, steps * dt, steps, 0, threads=8, reaction_nodes=top, reaction_coord=2)
    run.advance(steps)
    return m, run


def test_assembly_info_and_element_order(hip_device, monkeypatch):
    monkeypatch.setenv("NSM_TILE_ELEMS", "512")
    prob = N.CubeProblem(n=16, materials=[N.problems.STEEL, N.problems.ALUMINIUM], num_steps=1)
    mixed = prob.make_model("mixed")
    info = mixed.AssemblyInfo()
    assert info["tiled"] and info["num_tiles"] == 8 and mixed.TileElems() == 512
    assert info["interior_nodes"] + info["boundary_nodes"] == prob.num_nodes
    order = mixed.ElementOrder()
    assert np.array_equal(np.sort(order), np.arange(prob.num_elem))
    half = prob.num_elem // 2  # tiles keep elements inside their block
    assert order[:half].max() < half and order[half:].min() >= half
    # device node order: each tile's interior nodes, then the shared nodes;
    # host copies stay in mesh order
    node_order = mixed.NodeOrder()
    assert np.array_equal(np.sort(node_order), np.arange(prob.num_nodes))
    rng = np.random.default_rng(0)
    u = rng.normal(size=(prob.num_nodes, 3))
    mixed.SetVectorNodeData("displacement", u)
    assert np.array_equal(mixed.GetVectorNodeData("displacement"), u)
    assert np.array_equal(mixed.GetVectorNodeData("reference_coordinate"), prob.X)
    fp64 = prob.make_model("fp64")
    assert not fp64.AssemblyInfo()["tiled"]
    assert np.array_equal(fp64.ElementOrder(), np.arange(prob.num_elem))
    assert np.array_equal(fp64.NodeOrder(), np.arange(prob.num_nodes))


def test_tiled_internal_force_against_oracle_and_slots(hip_device, monkeypatch):
    prob = N.CubeProblem(n=12, materials=[N.problems.STEEL, N.problems.ALUMINIUM], num_steps=1, seed=11)
    rng = np.random.default_rng(4)
    u = rng.normal(0, 1e-3, (prob.num_nodes, 3))
    out = {}
    for mode in ("tiles", "slots"):
        monkeypatch.setenv("NSM_MIXED_ASSEMBLY", mode)
        m = prob.make_model("mixed")
        assert m.AssemblyInfo()["tiled"] == (mode == "tiles")
        m.SetVectorNodeData("displacement", u)
        m.ComputeInternalForce()
        out[mode] = (m.GetVectorNodeData("internal_force"), m.GetIntegrationPointData("deformation_gradient"),
                     m.GetIntegrationPointData("stress"))
        m.close()
    f, Fip, sip = oracle_model_from_problem(prob).internal_force(u, want_ip=True)
    ft, Ft, st = out["tiles"]
    print(f"tiled vs oracle: f {rel_err(ft, f):.2e} F {rel_err(Ft, Fip):.2e} sigma {rel_err(st, sip):.2e}")
    assert rel_err(ft, f) <= 1e-5 and rel_err(Ft, Fip) <= 1e-6 and rel_err(st, sip) <= 1e-5
    fs, Fs, ss = out["slots"]
    print(f"tiled vs slots: f {rel_err(ft, fs):.2e}")
    assert rel_err(ft, fs) <= 1e-6
    assert rel_err(Ft, Fs) <= 1e-6 and rel_err(st, ss) <= 1e-6


def test_tiled_run_on_renumbered_shuffled_mesh(hip_device):
    """Random node numbering and element order (tiles are then built from
    scattered ids); 60 steps against the oracle."""
    n = 10
    X, conn = N.jittered_cube(n, 21, 0.1)
    rng = np.random.default_rng(8)
    new_id = rng.permutation(len(X))
    Xn = np.zeros_like(X)
    Xn[new_id] = X
    conn = new_id[conn].astype(np.int32)
    half = n ** 3 // 2
    conn = np.concatenate([conn[:half][rng.permutation(half)], conn[half:][rng.permutation(n ** 3 - half)]])
    mats = [N.NeohookeanMaterial.from_young(7800, 200e9, 0.3), N.NeohookeanMaterial.from_young(2700, 70e9, 0.33)]
    steps = 60
    m, run = run_both(Xn, conn, [0, half, n ** 3], mats, steps)
    assert m.AssemblyInfo()["tiled"]
    eu = rel_err(m.GetVectorNodeData("displacement"), run.u)
    ev = rel_err(m.GetVectorNodeData("velocity"), run.v)
    ef = rel_err(m.GetVectorNodeData("internal_force"), run.f)
    eR = rel_err(m.GetReactionHistory(0, steps), run.reaction)
    print(f"shuffled mesh, tiled mixed vs oracle: u {eu:.2e} v {ev:.2e} f {ef:.2e} R {eR:.2e}")
    assert eu <= 1e-6 and ev <= 1e-5 and ef <= 1e-5 and eR <= 1e-5


def test_tiled_run_with_ten_blocks(hip_device):
    """More blocks than one element launch holds (8): two launch groups."""
    n = 10
    X, conn = N.jittered_cube(n, 5, 0.1)
    off = [k * n * n for k in range(n + 1)]  # one z-layer per block
    base = [(7800, 200e9, 0.3), (2700, 70e9, 0.33), (4500, 110e9, 0.34), (1200, 3e9, 0.45)]
    mats = [N.NeohookeanMaterial.from_young(*base[b % 4]) for b in range(n)]
    steps = 40
    m, run = run_both(X, conn, off, mats, steps)
    assert m.AssemblyInfo()["tiled"]
    eu = rel_err(m.GetVectorNodeData("displacement"), run.u)
    ef = rel_err(m.GetVectorNodeData("internal_force"), run.f)
    eR = rel_err(m.GetReactionHistory(0, steps), run.reaction)
    print(f"ten blocks, tiled mixed vs oracle: u {eu:.2e} f {ef:.2e} R {eR:.2e}")
    assert eu <= 1e-6 and ef <= 1e-5 and eR <= 1e-5
    # the last step evaluates F and sigma at the final displacement; returned in mesh order
    _, Fo, so = run.model.internal_force(run.u, want_ip=True)
    eF = rel_err(m.GetIntegrationPointData("deformation_gradient"), Fo)
    es = rel_err(m.GetIntegrationPointData("stress"), so)
    print(f"ten blocks: F {eF:.2e} sigma {es:.2e}")
    assert eF <= 1e-6 and es <= 1e-5


@pytest.mark.parametrize("tile_elems", [64, 128, 256])
def test_other_tile_sizes_against_oracle(hip_device, monkeypatch, tile_elems):
    """NSM_TILE_ELEMS=64 / 128 / 256 (1, 2 or 4 waves per tile; 512 is 8
    waves): tile sizes, a run and the tile-blocked F / sigma against
    the oracle."""
    monkeypatch.setenv("NSM_TILE_ELEMS", str(tile_elems))
    n = 10
    X, conn = N.jittered_cube(n, 9, 0.1)
    mats = [N.NeohookeanMaterial.from_young(7800, 200e9, 0.3), N.NeohookeanMaterial.from_young(2700, 70e9, 0.33)]
    steps = 40
    m, run = run_both(X, conn, [0, n ** 3 // 2, n ** 3], mats, steps)
    info = m.AssemblyInfo()
    assert info["tiled"] and info["num_tiles"] >= n ** 3 // tile_elems
    eu = rel_err(m.GetVectorNodeData("displacement"), run.u)
    ef = rel_err(m.GetVectorNodeData("internal_force"), run.f)
    eR = rel_err(m.GetReactionHistory(0, steps), run.reaction)
    _, Fo, so = run.model.internal_force(run.u, want_ip=True)
    eF = rel_err(m.GetIntegrationPointData("deformation_gradient"), Fo)
    es = rel_err(m.GetIntegrationPointData("stress"), so)
    print(f"{tile_elems}-element tiles vs oracle: u {eu:.2e} f {ef:.2e} R {eR:.2e} F {eF:.2e} sigma {es:.2e}")
    assert eu <= 1e-6 and ef <= 1e-5 and eR <= 1e-5 and eF <= 1e-6 and es <= 1e-5


def test_tiled_matches_slots_over_a_run(hip_device, monkeypatch):
    prob = N.CubeProblem(n=16, materials=[N.problems.STEEL, N.problems.ALUMINIUM], num_steps=200)
    res = {}
    for mode in ("tiles", "slots"):
        monkeypatch.setenv("NSM_MIXED_ASSEMBLY", mode)
        m = prob.make_model("mixed")
        integ = prob.make_integrator(m)
        integ.Integrate()
        res[mode] = (m.GetVectorNodeData("displacement"), m.GetReactionHistory(0, prob.num_steps))
        m.close()
    eu = rel_err(res["tiles"][0], res["slots"][0])
    eR = rel_err(res["tiles"][1], res["slots"][1])
    print(f"tiled vs slots after 200 steps: u {eu:.2e} R {eR:.2e}")
    assert eu <= 1e-7 and eR <= 1e-6


def test_gather_and_colour_assembly_agree(hip_device, monkeypatch):
    """Tiles of a lattice sum each node's forces per corner index; with
    NSM_TILE_ASSEMBLY=colors the same tiles sum them per element colour.  The
    FP64 sum of at most 8 FP32 forces is almost always exact in either order,
    so the two runs agree to 1e-12 (usually bit for bit)."""
    prob = N.CubeProblem(n=16, materials=[N.problems.STEEL, N.problems.ALUMINIUM], num_steps=200)
    res = {}
    for mode in ("gather", "colors"):
        if mode == "colors":
            monkeypatch.setenv("NSM_TILE_ASSEMBLY", "colors")
        m = prob.make_model("mixed")
        info = m.AssemblyInfo()
        assert info["tiled"] and info["gather_tiles"] == (info["num_tiles"] if mode == "gather" else 0)
        prob.make_integrator(m).Integrate()
        res[mode] = [m.GetVectorNodeData(f) for f in ("displacement", "velocity", "internal_force")]
        res[mode].append(m.GetReactionHistory(0, prob.num_steps))
        m.close()
    same = all(np.array_equal(a, b) for a, b in zip(res["gather"], res["colors"]))
    errs = [rel_err(a, b) for a, b in zip(res["gather"], res["colors"])]
    print("gather vs colours after 200 steps: u %.1e v %.1e f %.1e R %.1e" % tuple(errs), "bitwise" if same else "")
    assert max(errs) <= 1e-12


def test_mixed_gather_and_colour_tiles_against_oracle(hip_device):
    """Elements with their corners relabelled (a quarter turn about their z
    axis) hold nodes at the same corner index as a neighbour: their tiles use
    the colour order, the rest gather per corner, in one launch."""
    n = 12
    X, conn = N.jittered_cube(n, 13, 0.1)
    rot = np.arange(300, 900)
    conn[rot] = conn[rot][:, [1, 2, 3, 0, 5, 6, 7, 4]]
    mats = [N.NeohookeanMaterial.from_young(7800, 200e9, 0.3), N.NeohookeanMaterial.from_young(2700, 70e9, 0.33)]
    steps = 60
    m, run = run_both(X, conn, [0, n ** 3 // 2, n ** 3], mats, steps)
    info = m.AssemblyInfo()
    assert info["tiled"] and 0 < info["gather_tiles"] < info["num_tiles"]
    eu = rel_err(m.GetVectorNodeData("displacement"), run.u)
    ev = rel_err(m.GetVectorNodeData("velocity"), run.v)
    ef = rel_err(m.GetVectorNodeData("internal_force"), run.f)
    eR = rel_err(m.GetReactionHistory(0, steps), run.reaction)
    _, Fo, so = run.model.internal_force(run.u, want_ip=True)
    eF = rel_err(m.GetIntegrationPointData("deformation_gradient"), Fo)
    es = rel_err(m.GetIntegrationPointData("stress"), so)
    print(f"{info['gather_tiles']}/{info['num_tiles']} gather tiles vs oracle: u {eu:.2e} v {ev:.2e} f {ef:.2e} "
          f"R {eR:.2e} F {eF:.2e} sigma {es:.2e}")
    assert eu <= 1e-6 and ev <= 1e-5 and ef <= 1e-5 and eR <= 1e-5 and eF <= 1e-6 and es <= 1e-5


@pytest.mark.parametrize("mode", ["tiles", "slots", "fp64"])
def test_reaction_history_independent_of_call_chunking(hip_device, monkeypatch, mode):
    """A step's reaction total is summed inside the next step's first launch
    (tiled element kernel, or the node kernel of the slot and FP64 paths), or
    by a separate launch after a call's last step; both use the same fixed
    order, so chunking the run differently changes no bit."""
    if mode != "fp64":
        monkeypatch.setenv("NSM_MIXED_ASSEMBLY", mode)
    prob = N.CubeProblem(n=12, materials=[N.problems.STEEL, N.problems.ALUMINIUM], num_steps=40)
    hist = []
    for chunks in ([40], [1, 7, 32], [13, 1, 1, 25]):
        m = prob.make_model("fp64" if mode == "fp64" else "mixed")
        assert m.AssemblyInfo()["tiled"] == (mode == "tiles")
        integ = prob.make_integrator(m)
        for c in chunks:
            integ.Advance(c)
        hist.append((m.GetReactionHistory(0, 40), m.GetVectorNodeData("displacement")))
        m.close()
    for R, u in hist[1:]:
        assert np.array_equal(R, hist[0][0]) and np.array_equal(u, hist[0][1])
    assert np.all(hist[0][0][1:] != 0.0)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_nodes_without_elements(hip_device, precision):
    """Nodes no element touches have zero lumped mass and get the reference's
    a = (1/0)(0 + 0) = NaN (src/integrators/explicit_time_integrator.cc:252-255),
    exactly where the oracle has it; every other node is unaffected."""
    n = 6
    X, conn = N.jittered_cube(n, 13, 0.1)
    Xo = np.concatenate([X, [[0.3, 0.4, 0.5], [0.7, 0.2, 0.5], [0.5, 0.5, 0.5]]])
    mats = [N.NeohookeanMaterial.from_young(7800, 200e9, 0.3), N.NeohookeanMaterial.from_young(2700, 70e9, 0.33)]
    steps = 20
    m, run = run_both(Xo, conn, [0, n ** 3 // 2, n ** 3], mats, steps, precision)
    assert m.AssemblyInfo()["tiled"] == (precision == "mixed")
    u, v = m.GetVectorNodeData("displacement"), m.GetVectorNodeData("velocity")
    nan_ref = np.isnan(run.v).any(axis=1)
    assert nan_ref[-3:].all() and not nan_ref[:-3].any()
    assert np.array_equal(np.isnan(v).any(axis=1), nan_ref) and np.array_equal(np.isnan(u).any(axis=1), nan_ref)
    if precision == "fp64":
        assert np.array_equal(u, run.u, equal_nan=True) and np.array_equal(v, run.v, equal_nan=True)
    else:
        assert rel_err(u[:-3], run.u[:-3]) <= 1e-6 and rel_err(v[:-3], run.v[:-3]) <= 1e-5
    m.close()


def test_autotuned_tile_size(hip_device, monkeypatch, tmp_path):
    """Without NSM_TILE_ELEMS a small mixed model takes the autotuner's tile
    size (timed steps of each candidate), records it in the NSM_AUTOTUNE_CACHE
    file, and runs at it against the oracle; an ensemble reports what the
    autotuner timed."""
    monkeypatch.delenv("NSM_TILE_ELEMS", raising=False)
    prob = N.CubeProblem(n=9, materials=[N.problems.STEEL], num_steps=40)
    m = prob.make_model("mixed")
    te = m.TileElems()
    assert te in (64, 128, 256, 512)
    m2 = prob.make_model("mixed")  # same shape: the cached choice
    assert m2.TileElems() == te
    m2.close()
    integ = prob.make_integrator(m)
    integ.Integrate()
    om = oracle_model_from_problem(prob)
    run = O.ExplicitRun(om, prob.num_steps * prob.dt_factor * om.critical_dt(), prob.num_steps, 0,
                        reaction_nodes=prob.top, reaction_coord=2)
    run.advance(prob.num_steps)
    assert rel_err(m.GetVectorNodeData("displacement"), run.u) <= 1e-6
    assert rel_err(m.GetReactionHistory(0, prob.num_steps), run.reaction) <= 1e-5
    m.close()
    from nimblesm_amd.ensemble import EnsembleProblem, draw_parameters

    cache = tmp_path / "tiles.txt"
    monkeypatch.setenv("NSM_AUTOTUNE_CACHE", str(cache))
    ep = EnsembleProblem(draw_parameters(40, seed=5, n=(10, 12)), num_steps=20).build()
    t = ep.ens.AutotuneTimings()
    print(f"autotuner: model 9^3 -> {te}; ensemble {ep.ens.Info()['tile_elems']} from "
          + ", ".join(f"{k}: {1e6 * v:.1f} us" for k, v in t.items()), flush=True)
    assert sorted(t) == [64, 128, 256, 512] and all(v > 0 for v in t.values())
    assert ep.ens.Info()["tile_elems"] == min(t, key=t.get)
    assert cache.read_text().split()[-1] == str(ep.ens.Info()["tile_elems"])
    ep.close()


@pytest.mark.parametrize("precision,n", [("mixed", 12), ("fp64", 12), ("mixed", 4)])
def test_kernel_timing_leaves_results_unchanged(hip_device, precision, n):
    """Every 3rd launch of each kernel records its own start and stop events:
    the run's bits equal an untimed run's, and 4 of the 10 steps' element and
    node launches carry a duration.  The 4^3 cube is one tile with no
    boundary nodes, so its boundary launch has no work and its timer still
    completes."""
    mats = [N.problems.STEEL, N.problems.ALUMINIUM] if n > 4 else [N.problems.STEEL]
    prob = N.CubeProblem(n=n, materials=mats, num_steps=10).build()
    runs = []
    for every in (0, 3):
        m = prob.make_model(precision)
        if every:
            m.EnableKernelTiming(True, every=every)
        if n == 4:
            assert m.AssemblyInfo()["boundary_nodes"] == 0
        prob.make_integrator(m).Integrate()
        runs.append([m.GetVectorNodeData(f) for f in ("displacement", "velocity", "internal_force")])
        if every:
            el, ne = m.KernelTime("element")
            nd, nn = m.KernelTime("node")
            assert ne == 4 and nn == 4, (ne, nn)
            assert el > 0.0 and nd >= 0.0
            if n > 4 or precision == "fp64":
                assert nd > 0.0
        m.close()
    for a, b in zip(*runs):
        assert np.array_equal(a, b)
