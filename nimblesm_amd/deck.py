"""NimbleSM input decks, Genesis meshes and Exodus output for this path.

A NimbleSM run is described by a text deck (``key: value`` lines) naming a
Genesis mesh (Exodus II on NetCDF classic) and an Exodus results file.  This
module reads the deck keys the explicit Neo-Hookean path uses (reference
parser: Parser::ParseKeyValue, src/nimble_parser.cc:238-360; element blocks
BlockProperties, :159-169), reads the mesh's coordinates, element
blocks and node sets (as src/nimble_genesis_mesh.cc:63-317 does through the
Exodus II library), builds the device ModelData and integrator from them, and
writes nodal and element results as an Exodus II file at the output steps
(src/nimble_exodus_output.cc:79-292).  NetCDF classic is read and written
with scipy.io.netcdf_file; ids in the files are 1-based, everything here is
0-based.

    model, integ, mesh, deck = load_deck("wave_in_bar.in", precision="fp64")
    run_deck("wave_in_bar.in")           # runs and writes the deck's exodus output file
"""
from __future__ import annotations

import os
import re
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
from scipy.io import netcdf_file

from .model import Block, ExplicitTimeIntegrator, ModelData, NeohookeanMaterial

BC_KINDS = ("initial_velocity", "prescribed_velocity", "prescribed_displacement")
COORDS = {"x": 0, "y": 1, "z": 2}
LEN_STRING = 33  # Exodus II default name length + 1


# ----------------------------------------------------------------- Genesis
@dataclass
class GenesisMesh:
    X: np.ndarray                         # [N, 3] float64
    blocks: Dict[int, np.ndarray]         # block id -> [E_b, 8] int32 (0-based node ids)
    node_sets: Dict[int, np.ndarray]      # node-set id -> [n] int32 (0-based)
    title: str = ""

    @property
    def num_nodes(self) -> int:
        return int(self.X.shape[0])


def _chars_to_str(row) -> str:
    return b"".join(bytes(c) for c in row).decode("latin1").split("\x00")[0].strip()


def read_genesis(path: str) -> GenesisMesh:
    """Coordinates, HEX8 element blocks and node sets of a Genesis file."""
    with netcdf_file(path, "r", mmap=False) as f:
        v = f.variables
        if "coordx" in v:
            X = np.stack([v[c][:].astype(np.float64) for c in ("coordx", "coordy", "coordz")], axis=1)
        else:  # single "coord" variable [num_dim, num_nodes]
            X = np.ascontiguousarray(v["coord"][:].astype(np.float64).T)
        if X.shape[1] != 3:
            raise ValueError(f"{path}: only 3-D meshes are supported")
        blocks = {}
        for i, bid in enumerate(v["eb_prop1"][:]):
            c = v[f"connect{i + 1}"]
            etype = c._attributes.get("elem_type", b"HEX8")
            etype = etype.decode() if isinstance(etype, bytes) else str(etype)
            if not etype.upper().startswith("HEX") or c.shape[1] != 8:
                raise ValueError(f"{path}: block {int(bid)} is {etype} with {c.shape[1]} nodes; only HEX8 is supported")
            blocks[int(bid)] = np.ascontiguousarray(c[:].astype(np.int64) - 1).astype(np.int32)
        node_sets = {}
        if "ns_prop1" in v:
            for i, nid in enumerate(v["ns_prop1"][:]):
                node_sets[int(nid)] = (v[f"node_ns{i + 1}"][:].astype(np.int64) - 1).astype(np.int32)
        title = getattr(f, "title", b"")
        title = title.decode("latin1") if isinstance(title, bytes) else str(title)
    return GenesisMesh(X, blocks, node_sets, title)


def _put_names(f, var, dim, names):
    v = f.createVariable(var, "c", (dim, "len_string"))
    arr = np.zeros((len(names), LEN_STRING), dtype="S1")
    for i, n in enumerate(names):
        b = n.encode("latin1")[: LEN_STRING - 1]
        arr[i, : len(b)] = np.frombuffer(b, dtype="S1")
    v[:] = arr


def _write_mesh(f, mesh: GenesisMesh, title: str):
    f.title = title.encode("latin1")
    f.api_version = np.float32(5.22)
    f.version = np.float32(5.22)
    f.floating_point_word_size = np.int32(8)
    f.file_size = np.int32(1)
    N = mesh.num_nodes
    ids = sorted(mesh.blocks)
    f.createDimension("time_step", None)  # the record dimension comes first in NetCDF classic
    f.createDimension("len_string", LEN_STRING)
    f.createDimension("len_line", 81)
    f.createDimension("four", 4)
    f.createDimension("num_dim", 3)
    f.createDimension("num_nodes", N)
    f.createDimension("num_elem", int(sum(len(mesh.blocks[b]) for b in ids)))
    f.createDimension("num_el_blk", len(ids))
    for c, name in enumerate(("coordx", "coordy", "coordz")):
        f.createVariable(name, "d", ("num_nodes",))[:] = mesh.X[:, c]
    _put_names(f, "coor_names", "num_dim", ["x", "y", "z"])
    f.createVariable("eb_status", "i", ("num_el_blk",))[:] = np.ones(len(ids), np.int32)
    prop = f.createVariable("eb_prop1", "i", ("num_el_blk",))
    prop[:] = np.array(ids, np.int32)
    prop.name = b"ID"
    for i, bid in enumerate(ids):
        conn = mesh.blocks[bid]
        f.createDimension(f"num_el_in_blk{i + 1}", len(conn))
        f.createDimension(f"num_nod_per_el{i + 1}", 8)
        c = f.createVariable(f"connect{i + 1}", "i", (f"num_el_in_blk{i + 1}", f"num_nod_per_el{i + 1}"))
        c[:] = conn.astype(np.int32) + 1
        c.elem_type = b"HEX8"
    ns_ids = sorted(mesh.node_sets)
    if ns_ids:
        f.createDimension("num_node_sets", len(ns_ids))
        f.createVariable("ns_status", "i", ("num_node_sets",))[:] = np.ones(len(ns_ids), np.int32)
        p = f.createVariable("ns_prop1", "i", ("num_node_sets",))
        p[:] = np.array(ns_ids, np.int32)
        p.name = b"ID"
        for i, nid in enumerate(ns_ids):
            nodes = mesh.node_sets[nid]
            f.createDimension(f"num_nod_ns{i + 1}", len(nodes))
            f.createVariable(f"node_ns{i + 1}", "i", (f"num_nod_ns{i + 1}",))[:] = nodes.astype(np.int32) + 1
    f.createVariable("time_whole", "d", ("time_step",))


def write_genesis(path: str, mesh: GenesisMesh, title: str = "nimblesm_amd") -> None:
    """Write a Genesis (Exodus II, NetCDF classic) mesh file."""
    with netcdf_file(path, "w", version=2) as f:
        _write_mesh(f, mesh, title)


# ---------------------------------------------------------------- the deck
@dataclass
class InputDeck:
    genesis_input_file: str = ""
    exodus_output_file: str = ""
    initial_time: float = 0.0
    final_time: float = 0.0
    number_of_load_steps: int = 0
    output_frequency: int = 1
    output_fields: List[str] = field(default_factory=list)
    materials: Dict[str, Dict[str, float]] = field(default_factory=dict)
    element_blocks: List[Tuple[int, str]] = field(default_factory=list)  # (block id, material name)
    boundary_conditions: List[Tuple[str, int, int, float]] = field(default_factory=list)  # kind, node set, coord, value
    uq_model: str = ""                                                  # "file <samples file>"
    uq_parameters: Dict[str, str] = field(default_factory=dict)  # material -> "<name> <range> ..."
    directory: str = "."


# keys whose values this path cannot honour: parsing them is an error rather
# than a silent difference from the reference's run
_UNSUPPORTED = {"contact", "contact dicing", "contact splitting", "contact backend", "contact visualization",
                "nonlinear solver relative tolerance", "nonlinear solver maximum iterations"}


def _trailing_int(name: str) -> int:
    m = re.search(r"(\d+)\s*$", name)
    if not m:
        raise ValueError(f"cannot take an id from '{name}' (expected e.g. block_1, nodelist_2)")
    return int(m.group(1))


def parse_input_deck(path: str) -> InputDeck:
    """Keys of the explicit Neo-Hookean path (Parser::ParseKeyValue,
    src/nimble_parser.cc:238-360)."""
    deck = InputDeck(directory=os.path.dirname(os.path.abspath(path)))
    for lineno, raw in enumerate(open(path), 1):
        line = raw.split("#")[0].strip()
        if not line or ":" not in line:
            continue
        key, value = (s.strip() for s in line.split(":", 1))
        where = f"{path}:{lineno}"
        if key == "genesis input file":
            deck.genesis_input_file = value
        elif key == "exodus output file":
            deck.exodus_output_file = value
        elif key == "initial time":
            deck.initial_time = float(value)
        elif key == "final time":
            deck.final_time = float(value)
        elif key == "number of load steps":
            deck.number_of_load_steps = int(value)
        elif key == "output frequency":
            deck.output_frequency = int(value)
        elif key == "output fields":
            deck.output_fields = value.split()
        elif key == "time integration scheme":
            if value != "explicit":
                raise ValueError(f"{where}: only the explicit scheme is implemented, not '{value}'")
        elif key == "material parameters":
            tok = value.split()
            if len(tok) < 2 or tok[1] != "neohookean":
                raise ValueError(f"{where}: only neohookean materials are implemented ({value})")
            if len(tok) % 2:
                raise ValueError(f"{where}: material parameters must be name/value pairs")
            deck.materials[tok[0]] = {tok[i]: float(tok[i + 1]) for i in range(2, len(tok), 2)}
        elif key == "element block":
            tok = value.split()
            if len(tok) < 2:
                raise ValueError(f"{where}: element block needs a block name and a material name")
            deck.element_blocks.append((_trailing_int(tok[0]), tok[1]))
        elif key == "boundary condition":
            tok = value.split()
            if len(tok) != 4 or tok[0] not in BC_KINDS or tok[2] not in COORDS:
                raise ValueError(f"{where}: boundary condition '{value}' is not "
                                 f"<{'|'.join(BC_KINDS)}> <node set> <x|y|z> <constant>")
            try:
                mag = float(tok[3])
            except ValueError:
                raise ValueError(f"{where}: only constant boundary-condition magnitudes are implemented") from None
            deck.boundary_conditions.append((tok[0], _trailing_int(tok[1]), COORDS[tok[2]], mag))
        elif key == "uq parameters":
            # src/nimble_parser.cc:347-351: material key, then the rest of the line
            tok = value.split(None, 1)
            deck.uq_parameters[tok[0]] = tok[1] if len(tok) > 1 else ""
        elif key == "uq model":
            deck.uq_model = value
        elif key in _UNSUPPORTED:
            raise ValueError(f"{where}: '{key}' is outside this path")
        # other keys (timing files, MPI decomposition) do not change the results
    if deck.number_of_load_steps <= 0:
        raise ValueError(f"{path}: 'number of load steps' is missing")
    return deck


def _material(params: Dict[str, float], name: str) -> NeohookeanMaterial:
    try:
        return NeohookeanMaterial(params["density"], params["bulk_modulus"], params["shear_modulus"])
    except KeyError as e:
        raise ValueError(f"material {name}: missing {e.args[0]}") from None


def model_from_deck(deck: InputDeck, mesh: GenesisMesh, precision: str = "fp64") -> ModelData:
    mat_of_block = dict(deck.element_blocks)
    blocks = []
    for bid in sorted(mesh.blocks):  # ascending block id, as the reference's block map
        if bid not in mat_of_block:
            raise ValueError(f"mesh block {bid} has no 'element block' line in the deck")
        mname = mat_of_block[bid]
        if mname not in deck.materials:
            raise ValueError(f"block {bid}: material '{mname}' is not defined")
        blocks.append(Block(bid, mesh.blocks[bid], _material(deck.materials[mname], mname)))
    model = ModelData(mesh.X, blocks, precision)
    for kind, ns, coord, mag in deck.boundary_conditions:
        if ns not in mesh.node_sets:
            raise ValueError(f"node set {ns} is not in the mesh")
        model.AddBoundaryCondition(kind, mesh.node_sets[ns], coord, mag)
    return model


def load_deck(path: str, precision: str = "fp64"):
    """(model, integrator, mesh, deck) ready for Initialize()/Integrate().
    A deck with 'uq model' and 'uq parameters' lines also gets its off-nominal
    ensemble (model.uq, a nimblesm_amd.uq.UqModel; the scaled parameters are
    listed in parameter_samples.dat beside the deck, as the reference writes
    them)."""
    from .uq import uq_from_deck

    deck = parse_input_deck(path)
    mesh = read_genesis(os.path.join(deck.directory, deck.genesis_input_file))
    model = model_from_deck(deck, mesh, precision)
    model.uq = uq_from_deck(deck, model, write_samples=True)
    integ = ExplicitTimeIntegrator(model, deck.final_time, deck.number_of_load_steps, deck.output_frequency,
                                   deck.initial_time)
    return model, integ, mesh, deck


# ------------------------------------------------------------------ Exodus
_NODAL = ("displacement", "velocity", "acceleration", "internal_force", "lumped_mass", "reference_coordinate")
# node vectors of the UQ samples (src/uq/nimble_uq.cc:281-300): exact_<q>_<k>,
# sample_<q>_<k>, k from 1
_UQ_FIELD = re.compile(r"^(exact|sample)_(displacement|velocity|force)_(\d+)$")
_UQ_LABEL = {"displacement": "displacement", "velocity": "velocity", "force": "internal_force"}
_F_COMP = ("xx", "yy", "zz", "xy", "yz", "zx", "yx", "zy", "xz")  # K_F order, src/nimble_utils.h:101-109
_S_COMP = ("xx", "yy", "zz", "xy", "yz", "zx")                      # K_S order


class ExodusWriter:
    """Exodus II results file: the mesh, then one record per output time of
    the requested nodal vectors/scalars and element fields.  Element fields
    are ``deformation_gradient`` / ``stress`` (mean over the 8 Gauss points)
    and ``iptNN_deformation_gradient`` / ``iptNN_stress`` (Gauss point NN)."""

    def __init__(self, path: str, mesh: GenesisMesh, fields: Sequence[str]):
        self.mesh = mesh
        self.nodal, self.elem = [], []
        for name in fields:
            base = re.sub(r"^ipt0[1-8]_", "", name)
            if name in _NODAL or _UQ_FIELD.match(name):
                self.nodal.append(name)
            elif base in ("deformation_gradient", "stress"):
                self.elem.append(name)
            else:
                raise ValueError(f"output field '{name}' is not available on this path")
        self.f = netcdf_file(path, "w", version=2)
        _write_mesh(self.f, mesh, "nimblesm_amd")
        self.node_names = []
        for name in self.nodal:
            self.node_names += [name] if name == "lumped_mass" else [f"{name}_{c}" for c in "xyz"]
        self.elem_names = []
        for name in self.elem:
            comps = _F_COMP if name.endswith("deformation_gradient") else _S_COMP
            self.elem_names += [f"{name}_{c}" for c in comps]
        f = self.f
        if self.node_names:
            f.createDimension("num_nod_var", len(self.node_names))
            _put_names(f, "name_nod_var", "num_nod_var", self.node_names)
            for i in range(len(self.node_names)):
                f.createVariable(f"vals_nod_var{i + 1}", "d", ("time_step", "num_nodes"))
        if self.elem_names:
            f.createDimension("num_elem_var", len(self.elem_names))
            _put_names(f, "name_elem_var", "num_elem_var", self.elem_names)
            for i in range(len(self.elem_names)):
                for b in range(len(mesh.blocks)):
                    f.createVariable(f"vals_elem_var{i + 1}eb{b + 1}", "d", ("time_step", f"num_el_in_blk{b + 1}"))
        self.step = 0

    def write(self, time: float, model: ModelData) -> None:
        f, k = self.f, self.step
        f.variables["time_whole"][k] = time
        col = 0
        for name in self.nodal:
            if name == "lumped_mass":
                f.variables[f"vals_nod_var{col + 1}"][k, :] = model.GetScalarNodeData(name)
                col += 1
            else:
                m = _UQ_FIELD.match(name)
                if m:
                    j = int(m.group(3)) - 1  # exact samples first, then approximate ones
                    v = model.GetUqNodeData(j if m.group(1) == "exact" else model.num_exact_samples + j,
                                            _UQ_LABEL[m.group(2)])
                else:
                    v = model.GetVectorNodeData(name)
                for c in range(3):
                    f.variables[f"vals_nod_var{col + 1}"][k, :] = v[:, c]
                    col += 1
        if self.elem:
            ip = {"deformation_gradient": model.GetIntegrationPointData("deformation_gradient"),
                  "stress": model.GetIntegrationPointData("stress")}
            ids = sorted(self.mesh.blocks)
            col = 0
            for name in self.elem:
                m = re.match(r"^ipt0([1-8])_(.*)$", name)
                base = m.group(2) if m else name
                vals = ip[base][:, int(m.group(1)) - 1, :] if m else ip[base].mean(axis=1)
                for c in range(vals.shape[1]):
                    e0 = 0
                    for b, bid in enumerate(ids):
                        e1 = e0 + len(self.mesh.blocks[bid])
                        f.variables[f"vals_elem_var{col + 1}eb{b + 1}"][k, :] = vals[e0:e1, c]
                        e0 = e1
                    col += 1
        self.step += 1
        f.flush()

    def close(self) -> None:
        self.f.close()


def run_deck(path: str, precision: str = "fp64", exodus_path: Optional[str] = None):
    """Run a deck end to end: initial output, then one Exodus record after
    every output step (the reference's output schedule,
    src/integrators/explicit_time_integrator.cc:150, :187-190, :266-275).
    Returns (model, integrator, path of the results file or None)."""
    model, integ, mesh, deck = load_deck(path, precision)
    out = exodus_path or (os.path.join(deck.directory, deck.exodus_output_file) if deck.exodus_output_file else None)
    writer = ExodusWriter(out, mesh, deck.output_fields or ["displacement", "velocity"]) if out else None
    integ.Initialize()
    if writer:
        writer.write(integ.time_current, model)
        integ.Integrate(lambda step, t: writer.write(t, model))
        writer.close()
    else:
        integ.Integrate()
    return model, integ, out
