"""The bench line's byte model (bench.py algorithmic_bytes) on the CPU: the
per-element-step figures of the entry-point model at 128^3 (mixed 667.6 B,
FP64 1147.6 B per element-step: E (32 + 120 s) + N 152 with s = 4 or 8
bytes per Gauss-point word, over E), and the split over launches
adds back up to the whole step (tiled: element launch + boundary launch;
slots: element launches + node launch)."""
import numpy as np
import pytest

import bench
import nimblesm_amd as N


@pytest.fixture(scope="module")
def cube128():
    return N.workload("cube128_2blk").build()


@pytest.mark.parametrize("s_bytes,per_elem", [(4, 667.6), (8, 1147.6)])
def test_entry_point_bytes_per_element(cube128, s_bytes, per_elem):
    prob = cube128
    assembly = {"tiled": False}
    _, step = bench.algorithmic_bytes(prob, s_bytes, assembly)
    assert prob.num_elem == 2097152 and prob.num_nodes == 2146689
    assert round(step / prob.num_elem, 1) == per_elem


def test_tiled_split_adds_up(cube128):
    prob = cube128
    n_int = 1481544  # tile-interior nodes of the 512-element tiles at 128^3
    n_bnd = prob.num_nodes - n_int
    per_launch, step = bench.algorithmic_bytes(prob, 4, {"tiled": True, "interior_nodes": n_int,
                                                         "boundary_nodes": n_bnd})
    assert len(per_launch) == 1
    assert per_launch[0] == 1330863472  # the element launch's figure in the roofline
    assert per_launch[0] + n_bnd * 104 == step


def test_slot_split_adds_up():
    prob = N.workload("cube16_fp64").build()
    per_launch, step = bench.algorithmic_bytes(prob, 8, {"tiled": False})
    # one block, one element launch reading X and u of every node once; the
    # node launch carries the rest
    assert len(per_launch) == 1
    assert per_launch[0] + prob.num_nodes * 104 == step
    assert per_launch[0] == prob.num_elem * (32 + 120 * 8) + len(np.unique(prob.conn)) * 48
