"""Batched ensembles: many independent samples advanced together on the
device (the nsm_ensemble_* entry points of include/nimblesm_amd.h).

Each sample is a copy of one of the ensemble's meshes with its own block
materials, its own magnitudes of the mesh's boundary conditions and its own
time step.  The reference's counterpart is the serial loop over the exact
samples of its UQ model (src/uq/nimble_uq_model_data.cc:145-248), where all
samples share one mesh, lumped mass and time step; here each sample's
trajectory is the one a ModelData holding that sample alone computes.

``ensemble_workload`` builds the synthetic ensemble of the benchmarks: cubes
of n^3 HEX8 elements (one seeded jittered mesh per edge count), bottom face
fixed, top face driven at a per-sample z velocity, per-sample Young's
modulus, Poisson ratio and density drawn lognormal around (200 GPa, 0.3,
7800 kg/m^3) with the ratio clipped to [0.2, 0.45], and the top-face
reaction history as the quantity of interest.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

from ._capi import BC_KINDS, FIELDS, NSM_FP64, NSM_MIXED, check, lib
from .model import NeohookeanMaterial
from .problems import jittered_cube

ENSEMBLE_KERNELS = {"element": 0, "node": 1}


class Ensemble:
    """Host handle of an nsm_ensemble (precision "fp64" or "mixed";
    tile_elems 0 lets the autotuner pick the element-tile size)."""

    def __init__(self, precision: str = "mixed", tile_elems: int = 0):
        self.precision = precision
        h = C.c_void_p()
        check(lib().nsm_ensemble_create({"fp64": NSM_FP64, "mixed": NSM_MIXED}[precision], int(tile_elems),
                                        C.byref(h)))
        self._h = h
        self.mesh_blocks: List[int] = []
        self.mesh_nodes: List[int] = []
        self.mesh_bcs: List[int] = []
        self.sample_mesh: List[int] = []

    # -- lifetime
    def close(self):
        if getattr(self, "_h", None):
            lib().nsm_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- description
    def AddMesh(self, X: np.ndarray, block_conn: Sequence[np.ndarray]) -> int:
        """A mesh: reference coordinates [N, 3] and the connectivity of each
        block in storage order ([E_b, 8] int32, 0-based)."""
        X = np.ascontiguousarray(X, dtype=np.float64)
        conn = np.ascontiguousarray(np.concatenate([np.asarray(c, np.int32).reshape(-1, 8) for c in block_conn]))
        ne = np.array([len(np.asarray(c).reshape(-1, 8)) for c in block_conn], np.int64)
        mid = C.c_int()
        check(lib().nsm_ensemble_add_mesh(self._h, X.shape[0], X.ctypes.data, len(ne), ne.ctypes.data,
                                          conn.ctypes.data, C.byref(mid)))
        self.mesh_blocks.append(len(ne))
        self.mesh_nodes.append(X.shape[0])
        self.mesh_bcs.append(0)
        return mid.value

    def AddBoundaryCondition(self, mesh_id: int, kind: str, nodes, coordinate: int):
        n = np.ascontiguousarray(nodes, dtype=np.int32)
        check(lib().nsm_ensemble_add_boundary_condition(self._h, int(mesh_id), BC_KINDS[kind], int(coordinate),
                                                        len(n), n.ctypes.data))
        self.mesh_bcs[mesh_id] += 1

    def SetReactionNodes(self, mesh_id: int, nodes, coordinate: int = 2):
        n = np.ascontiguousarray(nodes, dtype=np.int32)
        check(lib().nsm_ensemble_set_reaction_nodes(self._h, int(mesh_id), len(n), n.ctypes.data, int(coordinate)))

    def _params(self, mesh_id, density, bulk, shear, bc_magnitude):
        nb, nc = self.mesh_blocks[mesh_id], self.mesh_bcs[mesh_id]
        rho = np.ascontiguousarray(density, np.float64).reshape(-1, nb)
        K = np.ascontiguousarray(bulk, np.float64).reshape(-1, nb)
        G = np.ascontiguousarray(shear, np.float64).reshape(-1, nb)
        S = rho.shape[0]
        mag = np.ascontiguousarray(np.zeros((S, nc)) if bc_magnitude is None else bc_magnitude,
                                   np.float64).reshape(S, nc)
        assert K.shape == rho.shape and G.shape == rho.shape
        return S, rho, K, G, mag

    def AddSamples(self, mesh_id: int, density, bulk_modulus, shear_modulus, bc_magnitude=None) -> int:
        """Samples of mesh mesh_id: density, moduli [S, blocks]; condition
        magnitudes [S, conditions of the mesh]; returns the first index."""
        S, rho, K, G, mag = self._params(mesh_id, density, bulk_modulus, shear_modulus, bc_magnitude)
        first = C.c_int()
        check(lib().nsm_ensemble_add_samples(self._h, int(mesh_id), S, rho.ctypes.data, K.ctypes.data,
                                             G.ctypes.data, mag.ctypes.data, C.byref(first)))
        self.sample_mesh.extend([mesh_id] * S)
        return first.value

    def SetSamples(self, first: int, density, bulk_modulus, shear_modulus, bc_magnitude=None):
        S, rho, K, G, mag = self._params(self.sample_mesh[first], density, bulk_modulus, shear_modulus,
                                         bc_magnitude)
        check(lib().nsm_ensemble_set_samples(self._h, int(first), S, rho.ctypes.data, K.ctypes.data, G.ctypes.data,
                                             mag.ctypes.data))

    def Setup(self):
        check(lib().nsm_ensemble_setup(self._h))
        return self

    @property
    def num_samples(self) -> int:
        return len(self.sample_mesh)

    # -- time stepping
    def CriticalTimeSteps(self) -> np.ndarray:
        out = np.zeros(self.num_samples)
        check(lib().nsm_ensemble_get_critical_time_steps(self._h, out.ctypes.data))
        return out

    def Run(self, num_steps: int, final_time, output_frequency: int = 0, use_graph: bool = True):
        """Every sample from rest through num_steps central-difference steps
        with dt_user = final_time[s] / num_steps."""
        ft = np.ascontiguousarray(np.broadcast_to(np.asarray(final_time, np.float64), (self.num_samples,)))
        check(lib().nsm_ensemble_run(self._h, int(num_steps), int(output_frequency), ft.ctypes.data,
                                     int(bool(use_graph))))
        self.last_steps = int(num_steps)

    # -- results
    def SampleReactionHistory(self, num_points: int = 64) -> np.ndarray:
        out = np.zeros((self.num_samples, num_points))
        check(lib().nsm_ensemble_sample_reaction_history(self._h, int(num_points), out.ctypes.data))
        return out

    def GetReactionHistory(self, sample: int, first_step: int = 0, count: Optional[int] = None) -> np.ndarray:
        count = self.last_steps - first_step if count is None else count
        out = np.zeros(count)
        check(lib().nsm_ensemble_get_reaction_history(self._h, int(sample), int(first_step), int(count),
                                                      out.ctypes.data))
        return out

    def GetNodeData(self, sample: int, label: str) -> np.ndarray:
        n = self.mesh_nodes[self.sample_mesh[sample]]
        out = np.zeros(n if label == "lumped_mass" else (n, 3))
        check(lib().nsm_ensemble_get_field(self._h, int(sample), FIELDS[label], out.ctypes.data))
        return out

    def Info(self) -> Dict[str, int]:
        a, b, c, d, f = (C.c_int64() for _ in range(5))
        t = C.c_int()
        check(lib().nsm_ensemble_info(self._h, C.byref(a), C.byref(b), C.byref(c), C.byref(d), C.byref(t),
                                      C.byref(f)))
        return dict(num_samples=a.value, num_nodes=b.value, num_elem=c.value, work_items=d.value,
                    tile_elems=t.value, device_bytes=f.value)

    def AutotuneTimings(self) -> Dict[int, float]:
        """Seconds per step of each candidate tile size the autotuner timed at
        setup (empty when the choice was cached or given)."""
        cnt = C.c_int()
        te = np.zeros(8, np.int32)
        sec = np.zeros(8)
        check(lib().nsm_ensemble_autotune_timings(self._h, 8, C.byref(cnt), te.ctypes.data, sec.ctypes.data))
        return {int(te[k]): float(sec[k]) for k in range(min(cnt.value, 8))}

    def EnableKernelTiming(self, enable=True):
        check(lib().nsm_ensemble_enable_kernel_timing(self._h, int(enable)))

    def KernelTime(self, kernel: str):
        ms, n = C.c_double(), C.c_int64()
        check(lib().nsm_ensemble_kernel_time(self._h, ENSEMBLE_KERNELS[kernel], C.byref(ms), C.byref(n)))
        return ms.value, n.value


# ---------------------------------------------------------------------------
# Synthetic ensemble workload
# ---------------------------------------------------------------------------
NOMINAL = (200e9, 0.30, 7800.0)  # Young's modulus [Pa], Poisson ratio, density [kg/m^3]


@dataclass
class Draws:
    """Per-sample parameters: Young's modulus, Poisson ratio, density, top
    velocity and cube edge count."""

    E: np.ndarray
    nu: np.ndarray
    rho: np.ndarray
    v_top: np.ndarray
    n: np.ndarray

    def __len__(self):
        return len(self.E)

    def subset(self, idx) -> "Draws":
        return Draws(self.E[idx], self.nu[idx], self.rho[idx], self.v_top[idx], self.n[idx])

    def features(self) -> np.ndarray:
        """(E, nu, rho, v_top) per sample, [S, 4]."""
        return np.stack([self.E, self.nu, self.rho, self.v_top], axis=1)


def lognormal(rng, mean: float, cov: float, size: int) -> np.ndarray:
    """Lognormal draws with the given mean and coefficient of variation."""
    s2 = np.log1p(cov * cov)
    return rng.lognormal(np.log(mean) - 0.5 * s2, np.sqrt(s2), size)


def draw_parameters(count: int, seed: int, n=32, cov: float = 0.10, nu_clip=(0.2, 0.45),
                    v_range=(0.5, 2.0), nominal=NOMINAL) -> Draws:
    """count seeded draws: E, nu, rho lognormal with coefficient of variation
    cov around `nominal` (nu clipped), v_top uniform in v_range; edge count n
    (an int, or a sequence to draw uniformly from)."""
    rng = np.random.default_rng(seed)
    E = lognormal(rng, nominal[0], cov, count)
    nu = np.clip(lognormal(rng, nominal[1], cov, count), *nu_clip)
    rho = lognormal(rng, nominal[2], cov, count)
    v = rng.uniform(v_range[0], v_range[1], count)
    sizes = np.full(count, int(n)) if np.isscalar(n) else rng.choice(np.asarray(n, int), count)
    return Draws(E, nu, rho, v, sizes.astype(int))


def config2_draws(seed: int = 2024, n_main: int = 256, n_uneven: int = 64) -> Draws:
    """The batched-ensemble workload: n_main samples on the 32^3 cube and
    n_uneven on cubes with edge counts drawn uniformly from {16,24,32,40,48}."""
    a = draw_parameters(n_main, seed, n=32)
    b = draw_parameters(n_uneven, seed + 1, n=(16, 24, 32, 40, 48))
    return Draws(*(np.concatenate([getattr(a, k), getattr(b, k)]) for k in ("E", "nu", "rho", "v_top", "n")))


@dataclass
class EnsembleProblem:
    """An Ensemble over draws: one jittered cube mesh per edge count
    (bottom fixed, top v_z prescribed per sample, top-face reaction in z),
    samples grouped by mesh.  order[k] is the draw held by ensemble sample k."""

    draws: Draws
    num_steps: int = 400
    precision: str = "mixed"
    tile_elems: int = 0
    seed: int = 1234
    jitter: float = 0.1
    dt_factor: float = 0.5
    ens: Ensemble = field(default=None, repr=False)
    order: np.ndarray = field(default=None, repr=False)
    meshes: Dict[int, tuple] = field(default_factory=dict, repr=False)

    def build(self) -> "EnsembleProblem":
        d = self.draws
        self.ens = Ensemble(self.precision, self.tile_elems)
        order = []
        for n in sorted(set(int(x) for x in d.n)):
            X, conn = jittered_cube(n, self.seed, self.jitter)
            self.meshes[n] = (X, conn)
            mid = self.ens.AddMesh(X, [conn])
            n1 = n + 1
            bottom = np.arange(n1 * n1, dtype=np.int32)
            top = bottom + n * n1 * n1
            for c in range(3):
                self.ens.AddBoundaryCondition(mid, "prescribed_velocity", bottom, c)
            self.ens.AddBoundaryCondition(mid, "prescribed_velocity", top, 2)
            self.ens.SetReactionNodes(mid, top, 2)
            idx = np.nonzero(d.n == n)[0]
            mats = [NeohookeanMaterial.from_young(d.rho[i], d.E[i], d.nu[i]) for i in idx]
            mag = np.zeros((len(idx), 4))
            mag[:, 3] = d.v_top[idx]
            self.ens.AddSamples(mid, [m.density for m in mats], [m.bulk_modulus for m in mats],
                                [m.shear_modulus for m in mats], mag)
            order.extend(idx.tolist())
        self.order = np.array(order)
        self.ens.Setup()
        return self

    def final_times(self) -> np.ndarray:
        """t_final = num_steps * dt_factor * critical dt of each sample."""
        return self.num_steps * (self.dt_factor * self.ens.CriticalTimeSteps())

    def run(self, use_graph: bool = True) -> np.ndarray:
        """Runs every sample; returns the 64-point QoI in draw order [S, 64]."""
        self.ens.Run(self.num_steps, self.final_times(), 0, use_graph)
        return self.qoi()

    def qoi(self, num_points: int = 64) -> np.ndarray:
        q = self.ens.SampleReactionHistory(num_points)
        out = np.empty_like(q)
        out[self.order] = q
        return out

    def close(self):
        if self.ens is not None:
            self.ens.close()
