"""The oracle against the reference's own regression data.

* unit_tests/test_nimble_utils.cc:47-95 -- Invert_Full33 known answers
* test/dynamics/single_elem_complex_displacement/single_elem_native_neohookean
  -- one element, one step, FP64 F / sigma at 8 points and nodal forces
* the other test/dynamics decks -- multi-step runs (5..732 elements, up to
  10,000 steps, two blocks) compared at the gold output times.

The gold files were written by NimbleSM builds we cannot reproduce here (FMA,
Kokkos atomics), so agreement is not bit-exact; the bounds below are ~1e6
tighter than the reference's own exodiff tolerances (~1e-6 x max).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import load_golden
from tests.helpers import golden_oracle_model, output_steps

KEY = {"displacement": "u", "velocity": "v", "acceleration": "a", "internal_force": "f"}


def test_invert_full33_known_answers():
    # unit_tests/test_nimble_utils.cc:52-73
    m = np.zeros(9)
    m[[0, 3, 8]] = [1.0, 0.0, 0.0]  # XX XY XZ
    m[[6, 1, 4]] = [1.0, 1.0, 0.0]  # YX YY YZ
    m[[5, 7, 2]] = [1.0, 1.0, 1.0]  # ZX ZY ZZ
    ref = np.zeros(9)
    ref[[0, 3, 8]] = [1.0, 0.0, 0.0]
    ref[[6, 1, 4]] = [-1.0, 1.0, 0.0]
    ref[[5, 7, 2]] = [0.0, -1.0, 1.0]
    det, inv = O.invert_full33(m)
    assert np.array_equal(inv, ref) and det == 1.0
    # unit_tests/test_nimble_utils.cc:75-93
    m[[0, 3, 8]] = [3.0, 0.0, 2.0]
    m[[6, 1, 4]] = [2.0, 0.0, -2.0]
    m[[5, 7, 2]] = [0.0, 1.0, 1.0]
    ref[[0, 3, 8]] = [0.2, 0.2, 0.0]
    ref[[6, 1, 4]] = [-0.2, 0.3, 1.0]
    ref[[5, 7, 2]] = [0.2, -0.3, 0.0]
    det, inv = O.invert_full33(m)
    assert np.array_equal(inv, ref) and det == 10.0


def _run_deck(name):
    g = load_golden(name)
    model = golden_oracle_model(g)
    nls, of = int(g["num_load_steps"]), int(g["output_frequency"])
    run = O.ExplicitRun(model, float(g["t_final"]), nls, of, float(g["t_initial"]), keep_ip=True)
    results = []
    for s in output_steps(nls, of):
        run.advance(s + 1 - run.step)
        results.append(dict(t=run.time, u=run.u.copy(), v=run.v.copy(), a=run.a.copy(), f=run.f.copy(),
                            F=run.Fip.copy(), s=run.sip.copy(), m=run.m.copy()))
    return g, results


def _gold_vec(g, field, ti):
    keys = [f"gold_{field}_{c}" for c in "xyz"]
    if not all(k in g for k in keys):
        return None
    return np.stack([g[k][ti] for k in keys], axis=1)


def _gold_ip(g, kind, comps, ti):
    out = []
    for q in range(8):
        keys = [f"gold_ipt{q + 1:02d}_{kind}_{c}" for c in comps]
        if not all(k in g for k in keys):
            return None
        out.append(np.stack([g[k][ti] for k in keys], axis=1))
    return np.stack(out, axis=1)


F_COMPS = ("xx", "yy", "zz", "xy", "yz", "zx", "yx", "zy", "xz")  # K_F order
S_COMPS = ("xx", "yy", "zz", "xy", "yz", "zx")


def test_single_element_native_neohookean():
    g, res = _run_deck("single_elem_native_neohookean")
    r = res[-1]
    assert r["t"] == g["gold_time"][1]
    assert np.array_equal(r["u"], _gold_vec(g, "displacement", 1))
    assert np.array_equal(_gold_ip(g, "deformation_gradient", F_COMPS, 1), r["F"])
    gs = _gold_ip(g, "stress", S_COMPS, 1)
    assert np.max(np.abs(r["s"] - gs)) <= 1e-14 * np.max(np.abs(gs))
    gf = _gold_vec(g, "internal_force", 1)
    assert np.max(np.abs(r["f"] - gf)) <= 1e-15 * np.max(np.abs(gf))
    # lumped mass from the gold acceleration a = f/m: rho*V/8 for this element
    ga = _gold_vec(g, "acceleration", 1)
    assert np.max(np.abs(r["a"] - ga)) <= 1e-15 * np.max(np.abs(ga))


# (deck, {field: bound on max|err| / max_t max|gold|}) -- rigid_body_motion's forces
# are round-off noise (rigid translation) and are checked absolutely instead.
DECKS = [
    ("simple_deformation_modes", dict(displacement=1e-13, velocity=1e-13, internal_force=1e-12)),
    ("rigid_body_motion", dict(displacement=1e-13, velocity=1e-13)),
    ("wave_in_bar", dict(displacement=1e-11, velocity=1e-9)),
    ("brick_with_fibers", dict(displacement=1e-11, velocity=1e-10)),
    ("notched_plate_native_neohookean", dict(displacement=1e-10, velocity=1e-9, acceleration=1e-8)),
]


@pytest.mark.parametrize("name,bounds", DECKS, ids=[d[0] for d in DECKS])
def test_multistep_decks(name, bounds):
    g, res = _run_deck(name)
    assert len(res) == len(g["gold_time"]) - 1
    # errors are scaled by the field's maximum over all output times, the way
    # the reference's exodiff files set one absolute tolerance per variable
    scale = {f: max(np.max(np.abs(_gold_vec(g, f, ti))) for ti in range(1, len(res) + 1)) for f in bounds}
    for ti, r in enumerate(res, start=1):
        assert abs(r["t"] - g["gold_time"][ti]) <= 1e-15 * abs(g["gold_time"][ti])
        for field, bound in bounds.items():
            gv = _gold_vec(g, field, ti)
            assert gv is not None, field
            err = np.max(np.abs(r[KEY[field]] - gv))
            assert err <= bound * scale[field], (ti, field, err / scale[field])
        if "gold_lumped_mass" in g:
            assert np.array_equal(r["m"], g["gold_lumped_mass"][ti])
    if name == "rigid_body_motion":
        # rigid translation at 50 m/s: internal force is round-off only
        gf = _gold_vec(g, "internal_force", len(res))
        assert np.max(np.abs(res[-1]["f"])) < 1e-7 and np.max(np.abs(gf)) < 1e-7
