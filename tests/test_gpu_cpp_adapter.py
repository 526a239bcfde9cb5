"""The C++ adapter classes of include/nimblesm_amd.hpp (ModelData,
NeohookeanMaterial, ExplicitTimeIntegrator) drive one explicit run in a C++
program (tests/cpp/adapter_run, built by __graft_entry__.build()); its
displacement, velocity, internal force and reaction history must be
bit-identical to the same run through the Python host layer."""
import os
import subprocess

import numpy as np
import pytest

import nimblesm_amd as N
from tests.conftest import ROOT

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "tests", "cpp", "adapter_run")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_cpp_adapter_run_matches_python(hip_device, tmp_path, monkeypatch, precision):
    assert os.path.exists(EXE), "tests/cpp/adapter_run is missing: run __graft_entry__.build()"
    # one tile size for both processes (each would otherwise time its own)
    monkeypatch.setenv("NSM_TILE_ELEMS", "128")
    n, steps = 8, 30
    prob = N.CubeProblem(n=n, materials=[N.problems.STEEL, N.problems.ALUMINIUM], num_steps=steps)
    out = tmp_path / "run.bin"
    args = [EXE, str(n), precision, str(steps), repr(prob.dt_factor), str(out)]
    for b in prob.blocks():
        args += [repr(b.material.density), repr(b.material.bulk_modulus), repr(b.material.shear_modulus)]
    print(" ".join(args[1:5]), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    data = np.fromfile(out, dtype=np.float64)
    n3 = 3 * prob.num_nodes
    assert data.size == 3 * n3 + steps
    u, v, f, R = data[:n3], data[n3:2 * n3], data[2 * n3:3 * n3], data[3 * n3:]

    m = prob.make_model(precision)
    integ = prob.make_integrator(m)
    integ.Integrate()
    assert np.array_equal(u, m.GetVectorNodeData("displacement").ravel())
    assert np.array_equal(v, m.GetVectorNodeData("velocity").ravel())
    assert np.array_equal(f, m.GetVectorNodeData("internal_force").ravel())
    assert np.array_equal(R, m.GetReactionHistory(0, steps))
    m.close()
