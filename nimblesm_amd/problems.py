"""Synthetic jittered-cube problems of the benchmark workloads.

Unit cube [0,1]^3 meshed n x n x n with HEX8, interior nodes jittered by
U(-0.1, 0.1) * h (SplitMix64 keyed by seed), Neo-Hookean blocks layered in z,
bottom face fixed, top face driven at a prescribed z velocity from rest, and
dt = 0.5 x the reference's critical time step estimate.  The top-face
reaction-force history R_z(step) = -sum_top f_int,z is recorded every step.

Workloads (block materials as (E [Pa], nu, rho [kg/m^3])):
  cube16_fp64        n=16,  1 block  (200e9, 0.30, 7800), 1000 steps
  cube128_2blk       n=128, blocks split at z=0.5 (200e9,0.30,7800) /
                     (70e9,0.33,2700), 500 steps
  cube256_4blk       n=256, four z-layers (200e9,0.30,7800) (70e9,0.33,2700)
                     (110e9,0.34,4500) (3e9,0.45,1200), 200 steps
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Tuple

import numpy as np

from ._capi import check, lib
from .model import Block, ExplicitTimeIntegrator, ModelData, NeohookeanMaterial

STEEL = (200e9, 0.30, 7800.0)
ALUMINIUM = (70e9, 0.33, 2700.0)
TITANIUM = (110e9, 0.34, 4500.0)
POLYMER = (3e9, 0.45, 1200.0)


@dataclass
class CubeProblem:
    n: int
    materials: List[Tuple[float, float, float]]
    num_steps: int
    v_top: float = 1.0
    seed: int = 1234
    jitter: float = 0.1
    dt_factor: float = 0.5
    name: str = "cube"
    # filled by build()
    X: np.ndarray = field(default=None, repr=False)
    conn: np.ndarray = field(default=None, repr=False)

    @property
    def num_nodes(self):
        return (self.n + 1) ** 3

    @property
    def num_elem(self):
        return self.n ** 3

    def build(self):
        if self.X is None:
            self.X, self.conn = jittered_cube(self.n, self.seed, self.jitter)
        return self

    def node_layer(self, k: int) -> np.ndarray:
        n1 = self.n + 1
        return (np.arange(n1 * n1, dtype=np.int32) + k * n1 * n1).astype(np.int32)

    @property
    def bottom(self):
        return self.node_layer(0)

    @property
    def top(self):
        return self.node_layer(self.n)

    def block_ranges(self):
        """Element ranges per block: equal z-layers (elements are k-slowest)."""
        nb = len(self.materials)
        per_layer = self.n * self.n
        bounds = [(self.n * b) // nb for b in range(nb + 1)]
        return [(bounds[b] * per_layer, bounds[b + 1] * per_layer) for b in range(nb)]

    def blocks(self):
        self.build()
        out = []
        for b, ((e0, e1), (E, nu, rho)) in enumerate(zip(self.block_ranges(), self.materials)):
            out.append(Block(b + 1, self.conn[e0:e1], NeohookeanMaterial.from_young(rho, E, nu)))
        return out

    def bc_list(self):
        """(kind, nodes, coordinate, magnitude) in application order."""
        return [
            ("prescribed_velocity", self.bottom, 0, 0.0),
            ("prescribed_velocity", self.bottom, 1, 0.0),
            ("prescribed_velocity", self.bottom, 2, 0.0),
            ("prescribed_velocity", self.top, 2, self.v_top),
        ]

    def make_model(self, precision="fp64") -> ModelData:
        m = ModelData(self.build().X, self.blocks(), precision=precision)
        for kind, nodes, c, mag in self.bc_list():
            m.AddBoundaryCondition(kind, nodes, c, mag)
        m.SetReactionNodes(self.top, 2)
        return m

    def make_integrator(self, model: ModelData, num_steps=None, output_frequency=0) -> ExplicitTimeIntegrator:
        """Lumped mass + critical dt on the device, then t_final = steps * 0.5 dt_crit."""
        model.ComputeLumpedMass()
        dt = self.dt_factor * model.GetCriticalTimeStep()
        steps = self.num_steps if num_steps is None else num_steps
        integ = ExplicitTimeIntegrator(model, final_time=steps * dt, num_load_steps=steps,
                                       output_frequency=output_frequency)
        integ.critical_time_step = dt / self.dt_factor
        model.ApplyInitialConditions()
        model.ApplyKinematicConditions(0.0, 0.0)
        return integ


def jittered_cube(n: int, seed: int = 1234, jitter: float = 0.1):
    X = np.zeros(((n + 1) ** 3, 3))
    conn = np.zeros((n ** 3, 8), dtype=np.int32)
    check(lib().nsm_jittered_cube(n, C.c_uint64(seed), jitter, X.ctypes.data, conn.ctypes.data))
    return X, conn


WORKLOADS = {
    "cube16_fp64": dict(n=16, materials=[STEEL], num_steps=1000),
    "cube128_2blk": dict(n=128, materials=[STEEL, ALUMINIUM], num_steps=500),
    "cube256_4blk": dict(n=256, materials=[STEEL, ALUMINIUM, TITANIUM, POLYMER], num_steps=200),
    # one GPU's share of cube128_2blk split into 2 x 2 x 2 bricks: a 64^3 brick of one block
    "cube64_1blk": dict(n=64, materials=[STEEL], num_steps=500),
}


def workload(name: str, **over) -> CubeProblem:
    kw = dict(WORKLOADS[name])
    kw.update(over)
    return CubeProblem(name=name, **kw)
