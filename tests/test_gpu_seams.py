"""GPU element / material seams against the oracle (FP64, reference order).

The HIP kernels reproduce NimbleSM's operation order with -ffp-contract=off
and the platform libm's cube-root method, so every output is bit-exact."""
import numpy as np
import pytest

import nimblesm_amd as N
from oracle import oracle as O
from tests.helpers import random_elements

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def elems():
    return random_elements(4099, seed=21)  # not a multiple of the 32-element workgroup


def _dev(a, device):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to(device)


def test_deformation_gradients_bit_exact(hip_device, elems):
    import torch

    X, u = elems
    F = torch.zeros((len(X), 8, 9), dtype=torch.float64, device=hip_device)
    N.HexElement.ComputeDeformationGradients(_dev(X, hip_device), _dev(u, hip_device), F)
    torch.cuda.synchronize()
    assert np.array_equal(F.cpu().numpy(), O.deformation_gradients(X, u))


def test_stress_matches_oracle(hip_device, elems):
    import torch

    X, u = elems
    Fo = O.deformation_gradients(X, u)
    for K, G in [(166.67e9, 76.92e9), (1.6e12, 0.8e12), (10.0e9, 1.034e9)]:
        mat = N.NeohookeanMaterial(7800.0, K, G)
        s = torch.zeros((len(X), 8, 6), dtype=torch.float64, device=hip_device)
        mat.GetStress(_dev(Fo, hip_device), s)
        torch.cuda.synchronize()
        assert np.array_equal(s.cpu().numpy(), O.stress(K, G, Fo))


def test_identity_and_zero_sized(hip_device):
    import torch

    F = torch.zeros((3, 9), dtype=torch.float64, device=hip_device)
    F[:, :3] = 1.0
    s = torch.full((3, 6), 7.0, dtype=torch.float64, device=hip_device)
    N.NeohookeanMaterial(1.0, 1e9, 1e9).GetStress(F, s)
    torch.cuda.synchronize()
    assert torch.all(s == 0)
    e = torch.zeros((0, 8, 3), dtype=torch.float64, device=hip_device)
    N.HexElement.ComputeDeformationGradients(e, e, torch.zeros((0, 8, 9), dtype=torch.float64, device=hip_device))


def test_nodal_forces_bit_exact(hip_device, elems):
    import torch

    X, u = elems
    so = O.stress(1.6e11, 0.8e11, O.deformation_gradients(X, u))
    f = torch.zeros((len(X), 8, 3), dtype=torch.float64, device=hip_device)
    N.HexElement.ComputeNodalForces(_dev(X, hip_device), _dev(u, hip_device), _dev(so, hip_device), f)
    torch.cuda.synchronize()
    assert np.array_equal(f.cpu().numpy(), O.nodal_forces(X, u, so))


def test_lumped_mass_and_char_length_bit_exact(hip_device, elems):
    import torch

    X, u = elems
    m = torch.zeros((len(X), 8), dtype=torch.float64, device=hip_device)
    N.HexElement.ComputeLumpedMass(7800.0, _dev(X, hip_device), m)
    cl = torch.zeros(len(X), dtype=torch.float64, device=hip_device)
    x = X + u
    N.HexElement.ComputeCharacteristicLength(_dev(x, hip_device), cl)
    torch.cuda.synchronize()
    mo = np.array([O.lumped_mass_elem(7800.0, X[e]) for e in range(len(X))])
    assert np.array_equal(m.cpu().numpy(), mo)
    assert np.array_equal(cl.cpu().numpy(), np.array([O.char_length(x[e]) for e in range(len(x))]))


def test_inverted_element_reports_error(hip_device):
    import torch

    X = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    u = np.zeros_like(X)
    u[4:, 2] = -2.0  # top face pushed below the bottom face: det J < 0
    F = torch.zeros((1, 8, 9), dtype=torch.float64, device=hip_device)
    with pytest.raises(N.NimbleError) as ei:
        N.HexElement.ComputeNodalForces(_dev(X[None], hip_device), _dev(u[None], hip_device),
                                        torch.zeros((1, 8, 6), dtype=torch.float64, device=hip_device),
                                        torch.zeros((1, 8, 3), dtype=torch.float64, device=hip_device))
    assert ei.value.status == 3
    N.HexElement.ComputeDeformationGradients(_dev(X[None], hip_device), _dev(np.zeros_like(u)[None], hip_device), F)


def test_stress_strided_layouts(hip_device, elems):
    """The strided stress seam on LayoutRight [E][8][9] and LayoutLeft
    (element-fastest) arrays, and the grouped one on arrays in groups of 8
    elements from an element that does not start a group, give the oracle's
    stress bit for bit."""
    import torch

    X, u = elems
    Fo = O.deformation_gradients(X, u)
    E = len(X)
    K, G = 166.67e9, 76.92e9
    ref = O.stress(K, G, Fo)
    mat = N.NeohookeanMaterial(7800.0, K, G)
    # LayoutRight: (e, q, c) at 72 e + 9 q + c
    Fr = _dev(Fo, hip_device)
    Sr = torch.zeros((E, 8, 6), dtype=torch.float64, device=hip_device)
    mat.GetStressStrided(E, 8, Fr.data_ptr(), (72, 9, 1), Sr.data_ptr(), (48, 6, 1))
    # LayoutLeft: (e, q, c) at e + E q + 8 E c
    Fl = _dev(np.ascontiguousarray(Fo.transpose(2, 1, 0)), hip_device)
    Sl = torch.zeros((6, 8, E), dtype=torch.float64, device=hip_device)
    mat.GetStressStrided(E, 8, Fl.data_ptr(), (1, E, 8 * E), Sl.data_ptr(), (1, E, 8 * E))
    # groups of 8 elements (an FP64 model's layout), elements [3, E): value
    # (e, q, c) at 64 nc (e / 8) + 64 c + 8 q + e % 8
    E8 = (E + 7) // 8 * 8
    Fg_h = np.zeros((E8 // 8, 9, 8, 8))
    Fg_h[np.arange(E) // 8, :, :, np.arange(E) % 8] = Fo.transpose(0, 2, 1)
    Fg = _dev(Fg_h, hip_device)
    Sg = torch.zeros((E8 // 8, 6, 8, 8), dtype=torch.float64, device=hip_device)
    mat.GetStressGrouped(3, E - 3, 8, Fg.data_ptr(), (576, 1, 8, 64), Sg.data_ptr(), (384, 1, 8, 64))
    torch.cuda.synchronize()
    assert np.array_equal(Sr.cpu().numpy(), ref)
    assert np.array_equal(Sl.cpu().numpy().transpose(2, 1, 0), ref)
    Sg_h = Sg.cpu().numpy()
    e = np.arange(3, E)
    assert np.array_equal(Sg_h[e // 8, :, :, e % 8].transpose(0, 2, 1), ref[3:])
    assert not Sg_h[0, :, :, :3].any()  # elements before first_elem untouched


def test_split_force_evaluation_matches_fused(hip_device):
    """ComputeInternalForce split around a stress plug-in (F_np1 on the
    device, the Neo-Hookean stress of each block through the grouped seam on
    the model's own arrays and stream, then the forces from that stress) is
    bit-identical to the fused evaluation, and to the oracle."""
    from tests.helpers import oracle_model_from_problem

    prob = N.CubeProblem(n=9, materials=[N.problems.STEEL, N.problems.ALUMINIUM, N.problems.POLYMER],
                         num_steps=1, seed=5)
    rng = np.random.default_rng(4)
    u = rng.normal(0, 2e-4, (prob.num_nodes, 3))
    fused = prob.make_model("fp64")
    fused.SetVectorNodeData("displacement", u)
    fused.ComputeInternalForce()
    split = prob.make_model("fp64")
    split.SetVectorNodeData("displacement", u)
    split.ComputeDeformationGradients()
    Fp, eb = split.DevicePointer("deformation_gradient")
    Sp, _ = split.DevicePointer("stress")
    assert eb == 8
    E = split.num_elem
    for b, blk in enumerate(split.blocks):
        e0, e1 = int(split.block_offsets[b]), int(split.block_offsets[b + 1])
        blk.material.GetStressGrouped(e0, e1 - e0, 8, Fp, (576, 1, 8, 64), Sp, (384, 1, 8, 64),
                                      stream=split.Stream())
    split.ComputeInternalForceFromStress()
    om = oracle_model_from_problem(prob)
    f, Fip, sip = om.internal_force(u, want_ip=True)
    for m in (fused, split):
        assert np.array_equal(m.GetVectorNodeData("internal_force"), f)
        assert np.array_equal(m.GetIntegrationPointData("deformation_gradient"), Fip)
        assert np.array_equal(m.GetIntegrationPointData("stress"), sip)
    mixed = prob.make_model("mixed")
    with pytest.raises(N.NimbleError) as ei:
        mixed.ComputeDeformationGradients()
    assert ei.value.status == 4
    for m in (fused, split, mixed):
        m.close()


@pytest.mark.parametrize("scale", [2.0 ** -110, 2.0 ** -90, 2.0 ** 90, 2.0 ** 110])
def test_quotients_at_extreme_exponents_bit_exact(hip_device, elems, scale):
    """The device shares one reciprocal among the quotients over one
    denominator (the inverses' determinants, the stress's J) when every
    operand's exponent lies in [-299, 300], and divides one by one
    otherwise.  Scaled deformation gradients (det F and J ~ scale^3),
    scaled elements (det a ~ scale^3) and extreme moduli put operands on both
    sides of that range; every output must still equal the oracle's IEEE
    divisions bit for bit.  The reference's eigen solver (absolute
    tolerances) returns NaN for most far-from-unit F, so the scaled-F check
    keeps the elements whose stress is finite."""
    import torch

    X, u = elems[0][:700], elems[1][:700]
    F0 = O.deformation_gradients(X, u)
    Fs = F0 * scale
    ref = O.stress(166.67e9, 76.92e9, Fs)
    keep = np.isfinite(ref).all(axis=(1, 2))
    assert keep.sum() >= 100
    s = torch.zeros((int(keep.sum()), 8, 6), dtype=torch.float64, device=hip_device)
    N.NeohookeanMaterial(7800.0, 166.67e9, 76.92e9).GetStress(_dev(Fs[keep], hip_device), s)
    torch.cuda.synchronize()
    assert np.array_equal(s.cpu().numpy().view(np.uint64), ref[keep].view(np.uint64))
    for K, G in [(1e300, 1e299), (1e-300, 1e-301)]:
        s = torch.zeros((len(X), 8, 6), dtype=torch.float64, device=hip_device)
        N.NeohookeanMaterial(7800.0, K, G).GetStress(_dev(F0, hip_device), s)
        torch.cuda.synchronize()
        assert np.array_equal(s.cpu().numpy().view(np.uint64), O.stress(K, G, F0).view(np.uint64)), (K, G)
    Xs, us = X * scale, u * scale
    so = O.stress(1.6e11, 0.8e11, F0)
    f = torch.zeros((len(X), 8, 3), dtype=torch.float64, device=hip_device)
    N.HexElement.ComputeNodalForces(_dev(Xs, hip_device), _dev(us, hip_device), _dev(so, hip_device), f)
    torch.cuda.synchronize()
    assert np.array_equal(f.cpu().numpy().view(np.uint64), O.nodal_forces(Xs, us, so).view(np.uint64))
