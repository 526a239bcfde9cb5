"""ctypes access to the CPU oracle.

TEST INFRASTRUCTURE ONLY.  This module is imported by ``tests/``, by
``__graft_entry__.smoke()`` (as the checker) and by ``bench.py``'s
``cpu_baseline`` leg.  The product package ``nimblesm_amd`` never imports it.

``libnsm_oracle.so`` is the C restatement of NimbleSM's explicit internal-force
path (``nsm_oracle.c``); ``_ref/libnsm_ref.so`` wraps NimbleSM's own element /
material sources and exists only where the reference tree was present at build
time.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnsm_oracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libnsm_ref.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")

_lib = None
_ref = None


def build(ref: bool = True) -> None:
    """Compile the oracle (and, if the reference tree exists, oracle/_ref)."""
    import subprocess

    subprocess.run(["make", "-s", "-C", HERE, "libnsm_oracle.so"], check=True)
    if ref and os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


class _Mesh(C.Structure):
    _fields_ = [
        ("num_nodes", C.c_long),
        ("num_elem", C.c_long),
        ("num_blocks", C.c_int),
        ("X", C.c_void_p),
        ("conn", C.c_void_p),
        ("blk_off", C.c_void_p),
        ("rho", C.c_void_p),
        ("K", C.c_void_p),
        ("G", C.c_void_p),
    ]


class _Bcs(C.Structure):
    _fields_ = [
        ("count", C.c_long),
        ("node", C.c_void_p),
        ("coord", C.c_void_p),
        ("kind", C.c_void_p),
        ("value", C.c_void_p),
    ]


class _State(C.Structure):
    _fields_ = [
        ("u", C.c_void_p),
        ("v", C.c_void_p),
        ("a", C.c_void_p),
        ("f", C.c_void_p),
        ("m", C.c_void_p),
        ("Fip", C.c_void_p),
        ("sip", C.c_void_p),
        ("time_current", C.c_double),
    ]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build(ref=False)
        L = C.CDLL(LIB_PATH)
        L.orc_hex_tables.argtypes = [_dp, _dp, _dp]
        L.orc_invert3x3.argtypes = [_dp, _dp, C.POINTER(C.c_int)]
        L.orc_invert3x3.restype = C.c_double
        L.orc_invert_full33.argtypes = [_dp, _dp]
        L.orc_invert_full33.restype = C.c_double
        L.orc_cos_of_acos_div3.argtypes = [C.c_double]
        L.orc_cos_of_acos_div3.restype = C.c_double
        L.orc_polar_decomp.argtypes = [_dp, _dp, _dp]
        L.orc_eigen_sym33_nonunit.argtypes = [_dp, _dp, _dp, _dp, _dp]
        L.orc_neohookean_stress.argtypes = [C.c_double, C.c_double, _dp, _dp]
        L.orc_hex_char_length.argtypes = [_dp]
        L.orc_hex_char_length.restype = C.c_double
        L.orc_hex_lumped_mass.argtypes = [C.c_double, _dp, _dp]
        L.orc_batch_deformation_gradients.argtypes = [C.c_long, _dp, _dp, _dp]
        L.orc_batch_stress.argtypes = [C.c_long, C.c_double, C.c_double, _dp, _dp]
        L.orc_batch_nodal_forces.argtypes = [C.c_long, _dp, _dp, _dp, _dp]
        L.orc_model_lumped_mass.argtypes = [C.POINTER(_Mesh), _dp]
        L.orc_model_critical_dt.argtypes = [C.POINTER(_Mesh), C.c_void_p]
        L.orc_model_critical_dt.restype = C.c_double
        L.orc_model_internal_force.argtypes = [C.POINTER(_Mesh), _dp, _dp, C.c_void_p, C.c_void_p]
        L.orc_explicit_steps.argtypes = [
            C.POINTER(_Mesh), C.POINTER(_Bcs), C.POINTER(_State), C.c_int, C.c_int, C.c_int, C.c_int,
            C.c_double, C.c_void_p, C.c_long, C.c_int, C.c_void_p, C.c_int,
        ]
        L.orc_apply_initial_conditions.argtypes = [C.POINTER(_Bcs), _dp]
        L.orc_apply_kinematic_bc.argtypes = [C.POINTER(_Bcs), C.c_double, C.c_double, _dp, _dp]
        L.orc_splitmix64.argtypes = [C.c_ulonglong]
        L.orc_splitmix64.restype = C.c_ulonglong
        L.orc_jittered_cube.argtypes = [C.c_int, C.c_ulonglong, C.c_double, _dp, _ip]
        _lib = L
    return _lib


def ref_lib():
    """NimbleSM's own element/material code (None when not built here)."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_PATH):
            return None
        L = C.CDLL(REF_PATH)
        L.ref_deformation_gradients.argtypes = [C.c_long, _dp, _dp, _dp]
        L.ref_nodal_forces.argtypes = [C.c_long, _dp, _dp, _dp, _dp]
        L.ref_lumped_mass.argtypes = [C.c_long, C.c_double, _dp, _dp]
        L.ref_char_length.argtypes = [C.c_long, _dp, _dp]
        L.ref_stress.argtypes = [C.c_long, C.c_double, C.c_double, _dp, _dp]
        L.ref_polar_decomp.argtypes = [C.c_long, _dp, _dp, _dp]
        L.ref_invert_full33.argtypes = [_dp, _dp]
        L.ref_invert_full33.restype = C.c_double
        L.ref_invert3x3.argtypes = [_dp, _dp]
        L.ref_invert3x3.restype = C.c_double
        L.ref_cos_of_acos_div3.argtypes = [C.c_double]
        L.ref_cos_of_acos_div3.restype = C.c_double
        _ref = L
    return _ref


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


# ---------------------------------------------------------------- element level
def hex_tables():
    ip, N, dN = np.zeros(24), np.zeros(64), np.zeros(192)
    lib().orc_hex_tables(ip, N, dN)
    return ip.reshape(8, 3), N.reshape(8, 8), dN.reshape(8, 8, 3)


def invert_full33(m):
    inv = np.zeros(9)
    det = lib().orc_invert_full33(_c(m, np.float64), inv)
    return det, inv


def invert3x3(m):
    inv = np.zeros(9)
    st = C.c_int(0)
    det = lib().orc_invert3x3(_c(np.asarray(m).reshape(9), np.float64), inv, C.byref(st))
    return det, inv.reshape(3, 3), st.value


def polar_decomp(F):
    V, R = np.zeros(6), np.zeros(9)
    lib().orc_polar_decomp(_c(F, np.float64), V, R)
    return V, R


def deformation_gradients(Xg, ug):
    Xg, ug = _c(Xg, np.float64), _c(ug, np.float64)
    E = Xg.shape[0]
    F = np.zeros((E, 8, 9))
    lib().orc_batch_deformation_gradients(E, Xg, ug, F)
    return F


def stress(K, G, F):
    F = _c(F, np.float64)
    P = F.size // 9
    s = np.zeros(F.shape[:-1] + (6,))
    lib().orc_batch_stress(P, K, G, F, s)
    return s


def nodal_forces(Xg, ug, sig):
    Xg, ug, sig = _c(Xg, np.float64), _c(ug, np.float64), _c(sig, np.float64)
    E = Xg.shape[0]
    f = np.zeros((E, 8, 3))
    lib().orc_batch_nodal_forces(E, Xg, ug, sig, f)
    return f


def lumped_mass_elem(rho, X):
    m = np.zeros(8)
    lib().orc_hex_lumped_mass(rho, _c(X, np.float64).reshape(24), m)
    return m


def char_length(x):
    return lib().orc_hex_char_length(_c(x, np.float64).reshape(24))


# ---------------------------------------------------------------- mesh + model
def jittered_cube(n: int, seed: int = 1234, jitter: float = 0.1):
    """Uniform n^3 HEX8 cube with interior nodes jittered by U(-j,j)*h."""
    N = (n + 1) ** 3
    X = np.zeros((N, 3))
    conn = np.zeros((n ** 3, 8), dtype=np.int32)
    lib().orc_jittered_cube(n, seed, jitter, X, conn)
    return X, conn


@dataclass
class Model:
    """Mesh + materials + BCs, blocks concatenated in ascending block id."""

    X: np.ndarray
    conn: np.ndarray
    blk_off: np.ndarray
    rho: np.ndarray
    K: np.ndarray
    G: np.ndarray
    bc_node: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    bc_coord: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    bc_kind: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    bc_value: np.ndarray = field(default_factory=lambda: np.zeros(0))

    def __post_init__(self):
        self.X = _c(self.X, np.float64)
        self.conn = _c(self.conn, np.int32)
        self.blk_off = _c(self.blk_off, np.int64)
        self.rho = _c(self.rho, np.float64)
        self.K = _c(self.K, np.float64)
        self.G = _c(self.G, np.float64)
        self.bc_node = _c(self.bc_node, np.int32)
        self.bc_coord = _c(self.bc_coord, np.int32)
        self.bc_kind = _c(self.bc_kind, np.int32)
        self.bc_value = _c(self.bc_value, np.float64)
        self._mesh = _Mesh(
            self.X.shape[0], self.conn.shape[0], len(self.rho),
            self.X.ctypes.data, self.conn.ctypes.data, self.blk_off.ctypes.data,
            self.rho.ctypes.data, self.K.ctypes.data, self.G.ctypes.data,
        )
        self._bcs = _Bcs(
            len(self.bc_node), self.bc_node.ctypes.data, self.bc_coord.ctypes.data,
            self.bc_kind.ctypes.data, self.bc_value.ctypes.data,
        )

    @property
    def num_nodes(self):
        return self.X.shape[0]

    @property
    def num_elem(self):
        return self.conn.shape[0]

    def lumped_mass(self):
        m = np.zeros(self.num_nodes)
        lib().orc_model_lumped_mass(C.byref(self._mesh), m)
        return m

    def critical_dt(self, u=None):
        p = None if u is None else _c(u, np.float64).ctypes.data
        return lib().orc_model_critical_dt(C.byref(self._mesh), p)

    def internal_force(self, u, want_ip=False):
        u = _c(u, np.float64)
        f = np.zeros((self.num_nodes, 3))
        Fip = np.zeros((self.num_elem, 8, 9)) if want_ip else None
        sip = np.zeros((self.num_elem, 8, 6)) if want_ip else None
        st = lib().orc_model_internal_force(
            C.byref(self._mesh), u, f,
            None if Fip is None else Fip.ctypes.data, None if sip is None else sip.ctypes.data,
        )
        if st:
            raise FloatingPointError("non-positive element Jacobian")
        return (f, Fip, sip) if want_ip else f


class ExplicitRun:
    """Central-difference integration state on the oracle."""

    def __init__(self, model: Model, t_final: float, num_load_steps: int, output_frequency: int = 1,
                 t_initial: float = 0.0, keep_ip: bool = False, reaction_nodes=None, reaction_coord: int = 2,
                 threads: int = 1):
        self.model = model
        N, E = model.num_nodes, model.num_elem
        self.u = np.zeros((N, 3))
        self.v = np.zeros((N, 3))
        self.a = np.zeros((N, 3))
        self.f = np.zeros((N, 3))
        self.m = model.lumped_mass()
        self.Fip = np.zeros((E, 8, 9)) if keep_ip else None
        self.sip = np.zeros((E, 8, 6)) if keep_ip else None
        self.num_load_steps = num_load_steps
        self.output_frequency = output_frequency
        self.dt_user = (t_final - t_initial) / num_load_steps
        self.step = 0
        self.threads = threads
        self.reaction_nodes = None if reaction_nodes is None else _c(reaction_nodes, np.int32)
        self.reaction_coord = reaction_coord
        self.reaction = np.zeros(num_load_steps)
        self._state = _State(
            self.u.ctypes.data, self.v.ctypes.data, self.a.ctypes.data, self.f.ctypes.data, self.m.ctypes.data,
            None if self.Fip is None else self.Fip.ctypes.data, None if self.sip is None else self.sip.ctypes.data,
            t_initial,
        )
        # explicit_time_integrator.cc:146-147
        lib().orc_apply_initial_conditions(C.byref(model._bcs), self.v.reshape(-1))
        lib().orc_apply_kinematic_bc(C.byref(model._bcs), 0.0, 0.0, self.u.reshape(-1), self.v.reshape(-1))

    @property
    def time(self):
        return self._state.time_current

    def advance(self, nsteps: int):
        end = min(self.step + nsteps, self.num_load_steps)
        rn = self.reaction_nodes
        st = lib().orc_explicit_steps(
            C.byref(self.model._mesh), C.byref(self.model._bcs), C.byref(self._state), self.step, end,
            self.num_load_steps, self.output_frequency, self.dt_user,
            None if rn is None else rn.ctypes.data, 0 if rn is None else len(rn), self.reaction_coord,
            self.reaction.ctypes.data if rn is not None else None, self.threads,
        )
        self.step = end
        if st:
            raise FloatingPointError("non-positive element Jacobian")
        return self


def sample_history(R, num_samples=64):
    """The 64-point quantity of interest: history R[0..n-1] (value after each
    step) sampled at p_k = (k+1) n / num_samples steps with linear
    interpolation between integer steps j, where step j >= 1 holds R[j-1]
    and step 0 is 0 (at rest).  The rule is defined with the device sampler
    (the reference has no such output); restated here as the checker."""
    R = np.asarray(R, np.float64)
    n = len(R)
    full = np.concatenate([[0.0], R])  # value at integer step j
    out = np.empty(num_samples)
    for k in range(num_samples):
        p = (k + 1) * n / num_samples
        j = min(int(p), n)
        w = p - j
        r0 = full[j]
        r1 = full[j + 1] if j < n else r0
        out[k] = (1.0 - w) * r0 + w * r1
    return out


def _sample_bcs(model: Model) -> Model:
    """The model's mesh with only the conditions UQ samples take: initial and
    prescribed velocities (prescribed displacements set the nominal velocity
    only, src/nimble_boundary_condition_manager.h:184-215)."""
    keep = model.bc_kind != 2
    return Model(model.X, model.conn, model.blk_off, model.rho, model.K, model.G, model.bc_node[keep],
                 model.bc_coord[keep], model.bc_kind[keep], model.bc_value[keep])


class UqRun:
    """The reference's off-nominal ensemble (UQ) on the oracle.

    * nominal trajectory: ExplicitRun of the model;
    * exact sample k: ExplicitRun of the model with block moduli
      K_exact[k], G_exact[k] (Block::ComputeInternalForce with
      use_alternative_parameters, src/uq/nimble_uq_block.cc:57-164) and the
      samples' conditions; its update is UpdateWithNewVelocity /
      UpdateWithNewDisplacement (src/uq/nimble_uq_model_data.cc:222-248),
      the same arithmetic as the nominal one;
    * approximate sample s: force f_s = f_nom * c[s,0], then
      f_s += f_exact_k * c[s,k+1] for k ascending (UqModel::ApplyClosure,
      src/uq/nimble_uq.cc:327-346) after each step's forces, and the same
      half kick / drift / conditions as the other trajectories
      (src/integrators/explicit_time_integrator.cc:197-270).

    Trajectories are independent until the closure, so they are advanced one
    after another within each step.  All share the lumped mass.
    """

    def __init__(self, model: Model, K_exact, G_exact, coeff, t_final: float, num_load_steps: int,
                 output_frequency: int = 1, t_initial: float = 0.0, reaction_nodes=None, reaction_coord: int = 2):
        nb = len(model.rho)
        K_exact = np.asarray(K_exact, np.float64).reshape(-1, nb)
        G_exact = np.asarray(G_exact, np.float64).reshape(-1, nb)
        self.coeff = np.asarray(coeff, np.float64).reshape(-1, K_exact.shape[0] + 1)
        kw = dict(output_frequency=output_frequency, t_initial=t_initial)
        self.nominal = ExplicitRun(model, t_final, num_load_steps, reaction_nodes=reaction_nodes,
                                   reaction_coord=reaction_coord, **kw)
        self.samples_model = sm = _sample_bcs(model)
        self.exact = []
        for k in range(K_exact.shape[0]):
            mk = Model(sm.X, sm.conn, sm.blk_off, sm.rho, K_exact[k], G_exact[k], sm.bc_node, sm.bc_coord,
                       sm.bc_kind, sm.bc_value)
            self.exact.append(ExplicitRun(mk, t_final, num_load_steps, **kw))
        na, N = self.coeff.shape[0], model.num_nodes
        self.u = np.zeros((na, N, 3))
        self.v = np.zeros((na, N, 3))
        self.f = np.zeros((na, N, 3))
        self.m = self.nominal.m
        self.num_load_steps = num_load_steps
        self.output_frequency = output_frequency
        self.dt_user = self.nominal.dt_user
        self.time_current = t_initial
        self.step = 0
        for s in range(na):
            lib().orc_apply_initial_conditions(C.byref(sm._bcs), self.v[s].reshape(-1))
            lib().orc_apply_kinematic_bc(C.byref(sm._bcs), 0.0, 0.0, self.u[s].reshape(-1), self.v[s].reshape(-1))

    @property
    def num_samples(self):
        return len(self.exact) + self.coeff.shape[0]

    def sample(self, i: int):
        """(u, v, f) of sample i, exact samples first."""
        if i < len(self.exact):
            r = self.exact[i]
            return r.u, r.v, r.f
        s = i - len(self.exact)
        return self.u[s], self.v[s], self.f[s]

    def _bc(self, t_cur, t_prev, s):
        lib().orc_apply_kinematic_bc(C.byref(self.samples_model._bcs), t_cur, t_prev, self.u[s].reshape(-1),
                                     self.v[s].reshape(-1))

    def advance(self, nsteps: int):
        end = min(self.step + nsteps, self.num_load_steps)
        inv_m = (1.0 / self.m)[:, None]
        for step in range(self.step, end):
            out = self.output_frequency != 0 and (step % self.output_frequency == 0
                                                  or step == self.num_load_steps - 1)
            t_prev = self.time_current
            t_cur = t_prev + self.dt_user
            delta = t_cur - t_prev
            half = 0.5 * delta
            self.nominal.advance(1)
            for r in self.exact:
                r.advance(1)
            for s in range(self.coeff.shape[0]):
                u, v = self.u[s], self.v[s]
                # first half: a = (1/m)(f + f_ext), v += dt/2 a, BC, u += dt v, BC
                v += (half * 1.0) * (inv_m * (self.f[s] + 0.0))
                self._bc(t_cur, t_prev, s)
                u += (delta * 1.0) * v
                self._bc(t_cur, t_prev, s)
                c = self.coeff[s]
                f = self.nominal.f * c[0]
                for k, r in enumerate(self.exact):
                    f = f + r.f * c[k + 1]
                self.f[s] = f
                v += (half * 1.0) * (inv_m * (f + 0.0))
                if out:
                    self._bc(t_cur, t_prev, s)
            self.time_current = t_cur
        self.step = end
        return self
