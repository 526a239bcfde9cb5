"""Host-side mirror of NimbleSM's element, material and model-data interfaces
on top of the C ABI.  Method names follow the reference so that a NimbleSM
user finds the same calls:

* ``HexElement.ComputeDeformationGradients / ComputeNodalForces /
  ComputeLumpedMass / ComputeCharacteristicLength`` -- src/nimble_element.h
* ``NeohookeanMaterial.GetStress`` -- src/nimble_material.h:240-286
* ``ModelData`` -- src/nimble_model_data_base.h:87-345
* ``ExplicitTimeIntegrator.Integrate`` --
  src/integrators/explicit_time_integrator.cc:61-325

Element / material seams take device tensors (any object with
``data_ptr()``, e.g. torch tensors on ``cuda``); the model owns its device
memory and exchanges host numpy arrays.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Callable, Optional, Sequence

import numpy as np

from ._capi import BC_KINDS, FIELDS, KERNELS, NSM_FP64, NSM_MIXED, check, lib

_ELEM_WIDTH = {"deformation_gradient": 9, "stress": 6}


def _ptr(t) -> int:
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        return t.data_ptr()
    return t.ctypes.data


def _stream(stream) -> Optional[int]:
    if stream is None:
        return None
    return getattr(stream, "cuda_stream", stream)


class HexElement:
    """8-node hexahedron with the 2x2x2 Gauss rule (src/nimble_element.h:225)."""

    dim = 3
    num_nodes = 8
    num_int_pts = 8

    @staticmethod
    def ComputeDeformationGradients(X_gathered, u_gathered, F, stream=None):
        E = X_gathered.shape[0]
        check(lib().nsm_hex_compute_deformation_gradients(E, _ptr(X_gathered), _ptr(u_gathered), _ptr(F),
                                                          _stream(stream)))

    @staticmethod
    def ComputeNodalForces(X_gathered, u_gathered, sigma, f, stream=None):
        E = X_gathered.shape[0]
        check(lib().nsm_hex_compute_nodal_forces(E, _ptr(X_gathered), _ptr(u_gathered), _ptr(sigma), _ptr(f),
                                                 _stream(stream)))

    @staticmethod
    def ComputeLumpedMass(density, X_gathered, m, stream=None):
        E = X_gathered.shape[0]
        check(lib().nsm_hex_compute_lumped_mass(E, float(density), _ptr(X_gathered), _ptr(m), _stream(stream)))

    @staticmethod
    def ComputeCharacteristicLength(x_gathered, out, stream=None):
        E = x_gathered.shape[0]
        check(lib().nsm_hex_compute_characteristic_length(E, _ptr(x_gathered), _ptr(out), _stream(stream)))


class NeohookeanMaterial:
    """Neo-Hookean material (src/nimble_material.h:425-532)."""

    def __init__(self, density: float, bulk_modulus: float, shear_modulus: float):
        self.density = float(density)
        self.bulk_modulus = float(bulk_modulus)
        self.shear_modulus = float(shear_modulus)

    @classmethod
    def from_young(cls, density, youngs_modulus, poissons_ratio):
        E, nu = float(youngs_modulus), float(poissons_ratio)
        return cls(density, E / (3.0 * (1.0 - 2.0 * nu)), E / (2.0 * (1.0 + nu)))

    def GetDensity(self):
        return self.density

    def GetBulkModulus(self):
        return self.bulk_modulus

    def GetShearModulus(self):
        return self.shear_modulus

    def GetStress(self, deformation_gradient_np1, stress_np1, stream=None):
        """sigma_np1[p] from F_np1[p] for every point p (device tensors)."""
        P = deformation_gradient_np1.numel() // 9
        check(lib().nsm_neohookean_get_stress(P, self.bulk_modulus, self.shear_modulus,
                                              _ptr(deformation_gradient_np1), _ptr(stress_np1), _stream(stream)))

    def GetStressStrided(self, num_elem, num_ipt, F_ptr, F_strides, sigma_ptr, sigma_strides, stream=None):
        """The stress of num_elem x num_ipt points stored at (element, point,
        component) strides, e.g. Kokkos LayoutLeft / LayoutRight views
        (device addresses)."""
        check(lib().nsm_neohookean_get_stress_strided(int(num_elem), int(num_ipt), self.bulk_modulus,
                                                      self.shear_modulus, F_ptr, *map(int, F_strides), sigma_ptr,
                                                      *map(int, sigma_strides), _stream(stream)))

    def GetStressGrouped(self, first_elem, num_elem, num_ipt, F_ptr, F_strides, sigma_ptr, sigma_strides,
                         stream=None):
        """The same over arrays whose elements are stored in groups of 8, e.g.
        one block of an FP64 model's STEP_NP1 arrays: strides (group,
        element, point, component), elements [first_elem, first_elem +
        num_elem) (device addresses)."""
        fs = np.ascontiguousarray(F_strides, np.int64)
        ss = np.ascontiguousarray(sigma_strides, np.int64)
        assert fs.size == 4 and ss.size == 4
        check(lib().nsm_neohookean_get_stress_grouped(int(first_elem), int(num_elem), int(num_ipt), self.bulk_modulus,
                                                      self.shear_modulus, F_ptr, fs.ctypes.data, sigma_ptr,
                                                      ss.ctypes.data, _stream(stream)))


@dataclass
class Block:
    block_id: int
    conn: np.ndarray  # [E_b, 8] int32, 0-based
    material: NeohookeanMaterial


class ModelData:
    """Device-resident model data (nimble_kokkos::ModelData counterpart)."""

    def __init__(self, reference_coordinates: np.ndarray, blocks: Sequence[Block], precision: str = "fp64"):
        X = np.ascontiguousarray(reference_coordinates, dtype=np.float64)
        assert X.ndim == 2 and X.shape[1] == 3
        blocks = sorted(blocks, key=lambda b: b.block_id)  # std::map order
        self.blocks = blocks
        self.precision = precision
        conn = np.ascontiguousarray(np.concatenate([np.asarray(b.conn, np.int32).reshape(-1, 8) for b in blocks]))
        ids = np.array([b.block_id for b in blocks], np.int32)
        ne = np.array([len(b.conn) for b in blocks], np.int64)
        rho = np.array([b.material.density for b in blocks])
        K = np.array([b.material.bulk_modulus for b in blocks])
        G = np.array([b.material.shear_modulus for b in blocks])
        prec = {"fp64": NSM_FP64, "mixed": NSM_MIXED}[precision]
        h = C.c_void_p()
        check(lib().nsm_model_create(X.shape[0], X.ctypes.data, len(blocks), ids.ctypes.data, ne.ctypes.data,
                                     conn.ctypes.data, rho.ctypes.data, K.ctypes.data, G.ctypes.data, prec,
                                     C.byref(h)))
        self._h = h
        self.num_nodes = X.shape[0]
        self.num_elem = conn.shape[0]
        self.block_offsets = np.concatenate([[0], np.cumsum(ne)])
        self.num_exact_samples = self.num_approx_samples = 0

    # -- lifetime
    def close(self):
        if getattr(self, "_h", None):
            lib().nsm_model_destroy(self._h)
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

    # -- boundary conditions (src/nimble_boundary_condition_manager.h)
    def AddBoundaryCondition(self, kind: str, node_set, coordinate, magnitude: float):
        nodes = np.ascontiguousarray(node_set, dtype=np.int32)
        coord = {"x": 0, "y": 1, "z": 2}.get(coordinate, coordinate)
        check(lib().nsm_model_add_boundary_condition(self._h, BC_KINDS[kind], int(coord), float(magnitude),
                                                     len(nodes), nodes.ctypes.data))

    def ApplyInitialConditions(self):
        check(lib().nsm_model_apply_initial_conditions(self._h))

    def ApplyKinematicConditions(self, time_current: float, time_previous: float):
        check(lib().nsm_model_apply_kinematic_conditions(self._h, time_current, time_previous))

    # -- setup
    def ComputeLumpedMass(self):
        check(lib().nsm_model_compute_lumped_mass(self._h))

    def GetCriticalTimeStep(self) -> float:
        dt = C.c_double()
        check(lib().nsm_model_get_critical_time_step(self._h, C.byref(dt)))
        return dt.value

    # -- hot path
    def ComputeInternalForce(self, time_previous=0.0, time_current=0.0, is_output_step=False):
        check(lib().nsm_model_compute_internal_force(self._h, time_previous, time_current, int(is_output_step)))

    def UpdateStates(self):
        check(lib().nsm_model_update_states(self._h))

    # -- the force evaluation split around a stress plug-in
    #    (src/nimble_kokkos_model_data.cc:1196-1257; FP64 models)
    def ComputeDeformationGradients(self):
        check(lib().nsm_model_compute_deformation_gradients(self._h))

    def ComputeInternalForceFromStress(self):
        check(lib().nsm_model_compute_internal_force_from_stress(self._h))

    def Stream(self) -> int:
        """The model's HIP stream (hipStream_t as an integer)."""
        p = C.c_void_p()
        check(lib().nsm_model_get_stream(self._h, C.byref(p)))
        return p.value

    def ExplicitSteps(self, step_begin, step_end, num_load_steps, output_frequency, dt_user, time_current):
        t = C.c_double(time_current)
        check(lib().nsm_model_explicit_steps(self._h, step_begin, step_end, num_load_steps, output_frequency,
                                             dt_user, C.byref(t)))
        return t.value

    # -- fields
    def GetVectorNodeData(self, label: str) -> np.ndarray:
        out = np.zeros((self.num_nodes, 3))
        check(lib().nsm_model_get_field(self._h, FIELDS[label], out.ctypes.data))
        return out

    def GetScalarNodeData(self, label: str) -> np.ndarray:
        out = np.zeros(self.num_nodes)
        check(lib().nsm_model_get_field(self._h, FIELDS[label], out.ctypes.data))
        return out

    def SetVectorNodeData(self, label: str, values: np.ndarray):
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(self.num_nodes, 3)
        check(lib().nsm_model_set_field(self._h, FIELDS[label], v.ctypes.data))

    def SetScalarNodeData(self, label: str, values: np.ndarray):
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(self.num_nodes)
        check(lib().nsm_model_set_field(self._h, FIELDS[label], v.ctypes.data))

    def GetIntegrationPointData(self, label: str) -> np.ndarray:
        w = _ELEM_WIDTH[label]
        out = np.zeros((self.num_elem, 8, w))
        check(lib().nsm_model_get_field(self._h, FIELDS[label], out.ctypes.data))
        return out

    def DevicePointer(self, label: str):
        """(device address, bytes per word); node fields are in the order
        given by NodeOrder(), integration-point fields are [comp][ipt][E] in
        the order given by ElementOrder()."""
        p, eb = C.c_void_p(), C.c_int()
        check(lib().nsm_model_field_device_pointer(self._h, FIELDS[label], C.byref(p), C.byref(eb)))
        return p.value, eb.value

    def NodeOrder(self) -> np.ndarray:
        """Mesh node stored at each device node position."""
        out = np.zeros(self.num_nodes, dtype=np.int64)
        check(lib().nsm_model_get_node_order(self._h, out.ctypes.data))
        return out

    def ElementOrder(self) -> np.ndarray:
        """Mesh element stored at each device element position."""
        out = np.zeros(self.num_elem, dtype=np.int64)
        check(lib().nsm_model_get_element_order(self._h, out.ctypes.data))
        return out

    def AssemblyInfo(self):
        t, a, b, c, g = C.c_int(), C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        check(lib().nsm_model_assembly_info(self._h, C.byref(t), C.byref(a), C.byref(b), C.byref(c), C.byref(g)))
        return dict(tiled=bool(t.value), num_tiles=a.value, interior_nodes=b.value, boundary_nodes=c.value,
                    gather_tiles=g.value)

    def TileElems(self) -> int:
        """Elements per element tile of a tiled mixed model (0 otherwise)."""
        t = C.c_int()
        check(lib().nsm_model_tile_elems(self._h, C.byref(t)))
        return t.value

    # -- reaction-force history
    def SetReactionNodes(self, nodes, coordinate=2):
        n = np.ascontiguousarray(nodes, dtype=np.int32)
        check(lib().nsm_model_set_reaction_nodes(self._h, len(n), n.ctypes.data, int(coordinate)))

    def GetReactionHistory(self, first_step: int, count: int) -> np.ndarray:
        out = np.zeros(count)
        check(lib().nsm_model_get_reaction_history(self._h, first_step, count, out.ctypes.data))
        return out

    def SampleReactionHistory(self, num_steps: int, num_samples: int = 64) -> np.ndarray:
        """Reaction history of steps [0, num_steps) at num_samples equally
        spaced times (sample k at (k+1)*num_steps/num_samples steps, linear
        between steps; value 0 at rest before step 0)."""
        out = np.zeros(num_samples)
        check(lib().nsm_model_sample_reaction_history(self._h, int(num_steps), int(num_samples), out.ctypes.data))
        return out

    # -- off-nominal ensemble (src/uq/nimble_uq_model_data.cc:145-248)
    def SetUqSamples(self, bulk_modulus, shear_modulus, coefficients):
        """Exact samples with moduli bulk_modulus, shear_modulus [num_exact,
        num_blocks] (blocks in ascending block id) and approximate samples
        with closure coefficients [num_approx, num_exact + 1]; every sample
        starts at rest.  Empty arrays remove the ensemble."""
        nb = len(self.blocks)
        K = np.ascontiguousarray(bulk_modulus, dtype=np.float64).reshape(-1, nb)
        G = np.ascontiguousarray(shear_modulus, dtype=np.float64).reshape(-1, nb)
        if K.shape != G.shape:
            raise ValueError("bulk and shear moduli need the same shape")
        ne = K.shape[0]
        c = np.ascontiguousarray(coefficients, dtype=np.float64).reshape(-1, ne + 1)
        check(lib().nsm_model_set_uq_samples(self._h, ne, K.ctypes.data, G.ctypes.data, c.shape[0], c.ctypes.data))
        self.num_exact_samples, self.num_approx_samples = ne, c.shape[0]

    def GetUqNodeData(self, sample: int, label: str) -> np.ndarray:
        """Node vector `label` of UQ sample `sample` (exact samples first)."""
        out = np.zeros((self.num_nodes, 3))
        check(lib().nsm_model_get_uq_field(self._h, int(sample), FIELDS[label], out.ctypes.data))
        return out

    def SetUqNodeData(self, sample: int, label: str, values: np.ndarray):
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(self.num_nodes, 3)
        check(lib().nsm_model_set_uq_field(self._h, int(sample), FIELDS[label], v.ctypes.data))

    # -- instrumentation
    def Synchronize(self):
        check(lib().nsm_model_synchronize(self._h))

    def EnableKernelTiming(self, enable=True, every: int = 1):
        """HIP events on every `every`-th launch of each kernel."""
        check(lib().nsm_model_enable_kernel_timing(self._h, int(every) if enable else 0))

    def KernelTime(self, kernel: str):
        ms, n = C.c_double(), C.c_int64()
        check(lib().nsm_model_kernel_time(self._h, KERNELS[kernel], C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def Info(self):
        a, b, c, d = (C.c_int64() for _ in range(4))
        check(lib().nsm_model_info(self._h, C.byref(a), C.byref(b), C.byref(c), C.byref(d)))
        return dict(num_nodes=a.value, num_elem=b.value, csr_entries=c.value, device_bytes=d.value)


class ExplicitTimeIntegrator:
    """Central-difference driver (src/integrators/explicit_time_integrator.cc).

    ``Integrate`` runs all steps on the device; ``callback(step, time)`` is
    invoked after each output step (host copies of fields can then be read
    from the model)."""

    def __init__(self, model: ModelData, final_time: float, num_load_steps: int, output_frequency: int = 1,
                 initial_time: float = 0.0):
        if final_time < initial_time:
            raise ValueError(f"Final time: {final_time} is less than initial time: {initial_time}")
        self.model = model
        self.num_load_steps = int(num_load_steps)
        self.output_frequency = int(output_frequency)
        self.dt_user = (final_time - initial_time) / num_load_steps
        self.time_current = initial_time
        self.step = 0
        self.critical_time_step = None

    def Initialize(self):
        m = self.model
        m.ComputeLumpedMass()
        self.critical_time_step = m.GetCriticalTimeStep()
        m.ApplyInitialConditions()
        m.ApplyKinematicConditions(self.time_current, self.time_current)
        return self

    def is_output_step(self, step: int) -> bool:
        f = self.output_frequency
        return f != 0 and (step % f == 0 or step == self.num_load_steps - 1)

    def Advance(self, nsteps: int):
        end = min(self.step + nsteps, self.num_load_steps)
        self.time_current = self.model.ExplicitSteps(self.step, end, self.num_load_steps, self.output_frequency,
                                                     self.dt_user, self.time_current)
        self.step = end
        return self

    def Integrate(self, callback: Optional[Callable[[int, float], None]] = None):
        """Every remaining step; ``callback`` after each output step.  With
        output_frequency 0 no step is an output step (the reference then
        writes only its initial record), so the callback is not invoked."""
        if callback is None or self.output_frequency == 0:
            return self.Advance(self.num_load_steps - self.step)
        while self.step < self.num_load_steps:
            nxt = self.step
            # the last load step is always an output step, so this ends
            while nxt < self.num_load_steps - 1 and not self.is_output_step(nxt):
                nxt += 1
            self.Advance(nxt + 1 - self.step)
            callback(nxt, self.time_current)
        return self
