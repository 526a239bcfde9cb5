"""NimbleSM decks end to end through nimblesm_amd.deck on the device: the
meshes and deck data of the reference's regression problems (golden
fixtures) are written as a Genesis file and a NimbleSM input deck, run by
run_deck (FP64), and the Exodus results are compared with the reference's
gold files at every output time, with the bounds of the oracle's own golden
test (the FP64 device path is bit-identical to the oracle)."""
import os

import numpy as np
import pytest
from scipy.io import netcdf_file

from nimblesm_amd import deck as D
from tests.conftest import load_golden

pytestmark = pytest.mark.gpu

KINDS = ("initial_velocity", "prescribed_velocity", "prescribed_displacement")


def write_problem(g, d, output_fields=("displacement", "velocity")):
    """Genesis mesh + deck text from a golden fixture; one node set per
    boundary-condition line (consecutive entries with one kind, coordinate
    and value)."""
    blocks, mats = {}, []
    for b, bid in enumerate(g["blk_ids"]):
        e0, e1 = int(g["blk_off"][b]), int(g["blk_off"][b + 1])
        blocks[int(bid)] = g["conn"][e0:e1].astype(np.int32)
        mats.append(f"material parameters: mat_{bid} neohookean density {float(g['rho'][b])!r} "
                    f"bulk_modulus {float(g['K'][b])!r} shear_modulus {float(g['G'][b])!r}")
        mats.append(f"element block: block_{bid} mat_{bid}")
    lines, sets = [], {}
    key = list(zip(g["bc_kind"], g["bc_coord"], g["bc_value"]))
    i = 0
    while i < len(key):
        j = i
        while j < len(key) and key[j] == key[i]:
            j += 1
        sid = len(sets) + 1
        sets[sid] = g["bc_node"][i:j].astype(np.int32)
        kind, coord, val = key[i]
        lines.append(f"boundary condition: {KINDS[int(kind)]} nodelist_{sid} {'xyz'[int(coord)]} {float(val)!r}")
        i = j
    D.write_genesis(os.path.join(d, "mesh.g"), D.GenesisMesh(g["X"], blocks, sets))
    text = "\n".join([
        "genesis input file: mesh.g",
        "exodus output file: out.e",
        f"initial time: {float(g['t_initial'])!r}",
        f"final time: {float(g['t_final'])!r}",
        f"number of load steps: {int(g['num_load_steps'])}",
        f"output frequency: {int(g['output_frequency'])}",
        "output fields: " + " ".join(output_fields),
        *mats, *lines, ""])
    path = os.path.join(d, "problem.in")
    open(path, "w").write(text)
    return path


BOUNDS = {
    "simple_deformation_modes": dict(displacement=1e-13, velocity=1e-13),
    "wave_in_bar": dict(displacement=1e-11, velocity=1e-9),
    "brick_with_fibers": dict(displacement=1e-11, velocity=1e-10),
}


@pytest.mark.parametrize("name", sorted(BOUNDS))
def test_reference_deck_runs_through_the_deck_reader(hip_device, tmp_path, name):
    g = load_golden(name)
    path = write_problem(g, str(tmp_path))
    model, integ, out = D.run_deck(path, precision="fp64")
    model.close()
    with netcdf_file(out, "r", mmap=False) as f:
        times = f.variables["time_whole"][:].copy()
        names = [D._chars_to_str(r) for r in f.variables["name_nod_var"][:]]
        vals = {n: f.variables[f"vals_nod_var{i + 1}"][:].copy() for i, n in enumerate(names)}
    # initial state + one record per output step, as the gold file
    assert len(times) == len(g["gold_time"])
    assert np.allclose(times, g["gold_time"], rtol=1e-15, atol=0)
    # errors scaled by the field's maximum over components and output times,
    # as in the oracle's golden test
    for field, bound in BOUNDS[name].items():
        scale = max(np.max(np.abs(g[f"gold_{field}_{c}"])) for c in "xyz")
        for c in "xyz":
            err = np.max(np.abs(vals[f"{field}_{c}"] - g[f"gold_{field}_{c}"]))
            assert err <= bound * scale, (field, c, err / scale)
