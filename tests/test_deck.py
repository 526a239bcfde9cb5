"""Input decks, Genesis meshes and Exodus output (nimblesm_amd/deck.py) on
the CPU: NetCDF round trips, the deck keys of the explicit Neo-Hookean path,
errors for what the path cannot honour, and the Exodus record layout."""
import os

import numpy as np
import pytest
from scipy.io import netcdf_file

from nimblesm_amd import deck as D
from oracle import oracle as O


def cube_mesh(n=3):
    X, conn = O.jittered_cube(n, 7, 0.1)
    half = n * n * (n // 2)
    bottom = np.nonzero(X[:, 2] == 0.0)[0].astype(np.int32)
    top = np.nonzero(X[:, 2] == 1.0)[0].astype(np.int32)
    return D.GenesisMesh(X, {1: conn[:half].astype(np.int32), 3: conn[half:].astype(np.int32)}, {1: bottom, 2: top})


DECK = """# explicit two-block cube
genesis input file:      cube.g
exodus output file:      cube.e
initial time:            0.0
final time:              2.0e-6
number of load steps:    40
output frequency:        10
output fields:           displacement velocity stress ipt03_deformation_gradient
material parameters:     steel neohookean density 7800.0 bulk_modulus 1.6667e11 shear_modulus 7.692e10
material parameters:     alu neohookean density 2700.0 bulk_modulus 6.863e10 shear_modulus 2.632e10
element block:           block_1 steel
element block:           block_3 alu
boundary condition:      prescribed_velocity nodelist_1 x 0.0
boundary condition:      prescribed_velocity nodelist_1 y 0.0
boundary condition:      prescribed_velocity nodelist_1 z 0.0
boundary condition:      prescribed_velocity nodelist_2 z 1.0
boundary condition:      initial_velocity nodelist_2 x 0.5
time integration scheme: explicit
write timing data file:  false
"""


def test_genesis_round_trip(tmp_path):
    mesh = cube_mesh()
    path = str(tmp_path / "cube.g")
    D.write_genesis(path, mesh)
    back = D.read_genesis(path)
    assert np.array_equal(back.X, mesh.X)
    assert sorted(back.blocks) == [1, 3] and sorted(back.node_sets) == [1, 2]
    for b in mesh.blocks:
        assert np.array_equal(back.blocks[b], mesh.blocks[b])
    for s in mesh.node_sets:
        assert np.array_equal(back.node_sets[s], mesh.node_sets[s])
    # the file carries the Exodus II 1-based ids
    with netcdf_file(path, "r", mmap=False) as f:
        assert f.variables["connect1"][:].min() == 1
        assert list(f.variables["eb_prop1"][:]) == [1, 3]


def test_deck_keys(tmp_path):
    p = tmp_path / "cube.in"
    p.write_text(DECK)
    d = D.parse_input_deck(str(p))
    assert d.genesis_input_file == "cube.g" and d.exodus_output_file == "cube.e"
    assert d.final_time == 2.0e-6 and d.number_of_load_steps == 40 and d.output_frequency == 10
    assert d.element_blocks == [(1, "steel"), (3, "alu")]
    assert d.materials["alu"] == {"density": 2700.0, "bulk_modulus": 6.863e10, "shear_modulus": 2.632e10}
    assert d.boundary_conditions[3] == ("prescribed_velocity", 2, 2, 1.0)
    assert d.boundary_conditions[4] == ("initial_velocity", 2, 0, 0.5)
    assert d.output_fields == ["displacement", "velocity", "stress", "ipt03_deformation_gradient"]


@pytest.mark.parametrize("line,msg", [
    ("material parameters: m elastic density 1 bulk_modulus 1 shear_modulus 1", "neohookean"),
    ("time integration scheme: quasistatic", "explicit"),
    ("boundary condition: prescribed_velocity nodelist_1 x 1.0*t", "constant"),
    ("contact: on", "outside this path"),
])
def test_deck_rejects_what_the_path_cannot_honour(tmp_path, line, msg):
    p = tmp_path / "bad.in"
    p.write_text(DECK + line + "\n")
    with pytest.raises(ValueError, match=msg):
        D.parse_input_deck(str(p))


class FakeModel:
    """Stands in for ModelData: fields as functions of the record number."""

    def __init__(self, N, E):
        self.N, self.E, self.k = N, E, 0

    def GetVectorNodeData(self, label):
        return np.full((self.N, 3), 1.0 + self.k) * np.array([1, 2, 3]) + (label == "velocity")

    def GetScalarNodeData(self, label):
        return np.full(self.N, 5.0)

    def GetIntegrationPointData(self, label):
        w = 9 if label == "deformation_gradient" else 6
        return np.arange(self.E * 8 * w, dtype=np.float64).reshape(self.E, 8, w) + self.k


def test_exodus_records(tmp_path):
    mesh = cube_mesh()
    E = sum(len(c) for c in mesh.blocks.values())
    path = str(tmp_path / "out.e")
    w = D.ExodusWriter(path, mesh, ["displacement", "velocity", "stress", "ipt03_deformation_gradient"])
    fake = FakeModel(mesh.num_nodes, E)
    for k, t in enumerate((0.0, 1.0e-6, 2.0e-6)):
        fake.k = k
        w.write(t, fake)
    w.close()
    with netcdf_file(path, "r", mmap=False) as f:
        names = [D._chars_to_str(r) for r in f.variables["name_nod_var"][:]]
        assert names == ["displacement_x", "displacement_y", "displacement_z", "velocity_x", "velocity_y", "velocity_z"]
        assert list(f.variables["time_whole"][:]) == [0.0, 1.0e-6, 2.0e-6]
        assert np.all(f.variables["vals_nod_var3"][2, :] == 9.0)  # displacement_z at record 2 = 3 * (1 + 2)
        enames = [D._chars_to_str(r) for r in f.variables["name_elem_var"][:]]
        assert enames[:6] == ["stress_xx", "stress_yy", "stress_zz", "stress_xy", "stress_yz", "stress_zx"]
        assert enames[6] == "ipt03_deformation_gradient_xx" and len(enames) == 15
        # stress_xx of block 2 (second block) at record 1: mean over Gauss points
        ip = fake.GetIntegrationPointData("stress")  # k = 2 after the loop; rebuild record 1
        ip1 = ip - 1.0
        e0 = len(mesh.blocks[1])
        assert np.allclose(f.variables["vals_elem_var1eb2"][1, :], ip1[e0:, :, 0].mean(axis=1))
        # ipt03_deformation_gradient_xy (K_F slot 3) of block 1 at record 0
        F0 = fake.GetIntegrationPointData("deformation_gradient") - 2.0
        assert np.array_equal(f.variables["vals_elem_var10eb1"][0, :], F0[:e0, 2, 3])


class FakeUqModel(FakeModel):
    num_exact_samples = 2

    def GetUqNodeData(self, sample, label):
        return np.full((self.N, 3), 1000.0 * self.k + 10.0 * (sample + 1) + (label == "internal_force"))


def test_exodus_uq_sample_fields(tmp_path):
    """exact_<q>_<k> is sample k-1, sample_<q>_<k> is sample num_exact + k-1,
    one row per record"""
    mesh = cube_mesh()
    E = sum(len(c) for c in mesh.blocks.values())
    path = str(tmp_path / "out.e")
    fields = ["displacement", "exact_displacement_2", "sample_force_1", "sample_displacement_3"]
    w = D.ExodusWriter(path, mesh, fields)
    fake = FakeUqModel(mesh.num_nodes, E)
    for k in range(4):
        fake.k = k
        w.write(1.0e-6 * k, fake)
    w.close()
    with pytest.raises(ValueError, match="not available"):
        D.ExodusWriter(str(tmp_path / "bad.e"), mesh, ["sample_stress_1"])
    with netcdf_file(path, "r", mmap=False) as f:
        names = [D._chars_to_str(r) for r in f.variables["name_nod_var"][:]]
        assert names == [f"{q}_{c}" for q in fields for c in "xyz"]
        k = np.arange(4) * 1000.0
        col = {n: f.variables[f"vals_nod_var{i + 1}"][:, 0] for i, n in enumerate(names)}
        assert np.array_equal(col["exact_displacement_2_y"], k + 20.0)
        assert np.array_equal(col["sample_force_1_z"], k + 31.0)
        assert np.array_equal(col["sample_displacement_3_x"], k + 50.0)


REF_TEST = "/root/reference/test"
REF_DECKS = {  # fixture name -> (directory, mesh, deck) under the reference's test tree
    "single_elem_native_neohookean": ("dynamics/single_elem_complex_displacement", "single_elem.g",
                                      "single_elem_native_neohookean.in"),
    "simple_deformation_modes": ("dynamics/simple_deformation_modes", "simple_deformation_modes.g",
                                 "simple_deformation_modes.in"),
    "rigid_body_motion": ("dynamics/rigid_body_motion", "rigid_body_motion.g", "rigid_body_motion.in"),
    "wave_in_bar": ("dynamics/wave_in_bar", "wave_in_bar.g", "wave_in_bar.in"),
    "brick_with_fibers": ("dynamics/brick_with_fibers", "brick_with_fibers.g", "brick_with_fibers.in"),
    "notched_plate_native_neohookean": ("dynamics/notched_plate_native_neohookean",
                                        "notched_plate_native_neohookean.g", "notched_plate_native_neohookean.in"),
    "uq_wave": ("uq/uq_wave", "uq_wave.g", "uq_wave.in"),
}


@pytest.mark.skipif(not os.path.isdir(REF_TEST), reason="NimbleSM test tree not present")
@pytest.mark.parametrize("name", sorted(REF_DECKS))
def test_read_reference_genesis_meshes(name):
    """read_genesis on the reference's own meshes (src/nimble_genesis_mesh.cc:63-317
    reads them through Exodus II) and parse_input_deck on its decks agree
    with the committed fixtures, which a separate reader
    (tests/golden/make_golden.py) extracted: coordinates and connectivity bit
    for bit, blocks in ascending id with their element offsets, and the
    deck's conditions expanded over the mesh's node sets in deck order."""
    from tests.conftest import load_golden

    sub, mesh_name, deck_name = REF_DECKS[name]
    d = os.path.join(REF_TEST, sub)
    mesh = D.read_genesis(os.path.join(d, mesh_name))
    g = load_golden(name)
    assert np.array_equal(mesh.X, g["X"])
    ids = sorted(mesh.blocks)
    assert ids == [int(b) for b in g["blk_ids"]]
    off = np.concatenate([[0], np.cumsum([len(mesh.blocks[b]) for b in ids])])
    assert np.array_equal(off, g["blk_off"])
    assert np.array_equal(np.concatenate([mesh.blocks[b] for b in ids]), g["conn"])
    deck = D.parse_input_deck(os.path.join(d, deck_name))
    assert deck.genesis_input_file == mesh_name
    kinds = {k: i for i, k in enumerate(D.BC_KINDS)}
    bn, bc, bk, bv = [], [], [], []
    for kind, ns, coord, mag in deck.boundary_conditions:
        nodes = mesh.node_sets[ns]
        bn += list(nodes)
        bc += [coord] * len(nodes)
        bk += [kinds[kind]] * len(nodes)
        bv += [mag] * len(nodes)
    assert np.array_equal(np.array(bn, np.int32), g["bc_node"])
    assert np.array_equal(np.array(bc, np.int32), g["bc_coord"])
    assert np.array_equal(np.array(bk, np.int32), g["bc_kind"])
    assert np.array_equal(np.array(bv), g["bc_value"])
    assert deck.number_of_load_steps == int(g["num_load_steps"])
    assert deck.final_time == float(g["t_final"])
