"""Helpers shared by the test modules (inputs, error measures, oracle runs)."""
from __future__ import annotations

import numpy as np

from oracle import oracle as O

UNIT_HEX = np.array(
    [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], dtype=np.float64
)

# K_F storage index of row-major (i, j), src/nimble_utils.h:101-109
KF_OF = np.array([[0, 3, 8], [6, 1, 4], [5, 7, 2]])


def kf_to_mat(F):
    """[..., 9] K_F storage -> [..., 3, 3] matrices."""
    F = np.asarray(F)
    return F[..., KF_OF.reshape(-1)].reshape(F.shape[:-1] + (3, 3))


def mat_to_kf(A):
    A = np.asarray(A)
    out = np.zeros(A.shape[:-2] + (9,))
    for i in range(3):
        for j in range(3):
            out[..., KF_OF[i, j]] = A[..., i, j]
    return out


def sym_to_mat(s):
    s = np.asarray(s)
    idx = np.array([[0, 3, 5], [3, 1, 4], [5, 4, 2]])
    return s[..., idx.reshape(-1)].reshape(s.shape[:-1] + (3, 3))


def random_elements(E, seed=0, amps=(0.0, 1e-9, 1e-6, 1e-3, 0.03)):
    """Distorted hexes with displacements of mixed magnitudes; keeps only
    elements whose Gauss-point F has det > 0.2 everywhere."""
    rng = np.random.default_rng(seed)
    X = UNIT_HEX[None] * rng.uniform(0.5, 2.0, (E, 1, 3)) + rng.uniform(-0.1, 0.1, (E, 8, 3))
    u = rng.normal(0, 1, (E, 8, 3)) * np.asarray(amps)[rng.integers(0, len(amps), E)][:, None, None]
    F = O.deformation_gradients(X, u)
    det = np.linalg.det(kf_to_mat(F))
    ok = (det > 0.2).all(axis=1)
    return np.ascontiguousarray(X[ok]), np.ascontiguousarray(u[ok])


def rel_err(a, b):
    """Norm-wise relative error ||a-b|| / ||b|| (0 when both vanish)."""
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    d = np.linalg.norm(a - b)
    return d / nb if nb > 0 else d


def max_rel_err(a, b):
    """max |a-b| / max |b|."""
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    mb = np.max(np.abs(b)) if b.size else 0.0
    d = np.max(np.abs(a - b)) if b.size else 0.0
    return d / mb if mb > 0 else d


def oracle_model_from_problem(prob):
    """Oracle Model of a nimblesm_amd CubeProblem (same mesh, blocks, BCs)."""
    prob.build()
    blocks = prob.blocks()
    off = [0]
    rho, K, G = [], [], []
    for b in blocks:
        off.append(off[-1] + len(b.conn))
        rho.append(b.material.density)
        K.append(b.material.bulk_modulus)
        G.append(b.material.shear_modulus)
    kinds = {"initial_velocity": 0, "prescribed_velocity": 1, "prescribed_displacement": 2}
    bn, bc, bk, bv = [], [], [], []
    for kind, nodes, c, mag in prob.bc_list():
        for n in nodes:
            bn.append(n)
            bc.append(c)
            bk.append(kinds[kind])
            bv.append(mag)
    conn = np.concatenate([b.conn for b in blocks])
    return O.Model(prob.X, conn, np.array(off), np.array(rho), np.array(K), np.array(G), np.array(bn, np.int32),
                   np.array(bc, np.int32), np.array(bk, np.int32), np.array(bv))


def golden_oracle_model(g):
    return O.Model(g["X"], g["conn"], g["blk_off"], g["rho"], g["K"], g["G"], g["bc_node"], g["bc_coord"],
                   g["bc_kind"], g["bc_value"])


def output_steps(num_load_steps, output_frequency):
    return [s for s in range(num_load_steps) if output_frequency and (s % output_frequency == 0 or
                                                                      s == num_load_steps - 1)]
