"""Off-nominal ensemble (the reference's UQ model) on the device model.

NimbleSM's UQ build advances, beside the nominal solution, "exact" samples
that solve the same problem with other bulk / shear moduli and "approximate"
samples driven by a linear closure of the nominal and exact internal forces
(src/uq/nimble_uq.h:59-190, src/uq/nimble_uq.cc, src/uq/nimble_uq_model_data.cc).

``UqModel`` keeps the reference's host-side steps -- deck lines, samples file,
parameter scaling, the parameter listing -- and hands the scaled moduli and
closure coefficients to ``ModelData.SetUqSamples``; every trajectory then
runs on the device inside ``ExplicitSteps`` (exact samples through the
element kernels with their own moduli, approximate samples through one
closure + node-update kernel).

Samples file (UqModel::ReadSamples, src/uq/nimble_uq.cc:145-218)::

    <num exact> <num approx>
    <xi_1> ... <xi_P>                          one line per exact sample
    <xi_1> ... <xi_P> : <c_0> <c_1> ... <c_E>  one line per approximate sample

with P uncertain parameters and E exact samples; c_0 multiplies the nominal
force, c_k the force of exact sample k.  Parameter p of sample s is
``p0 + range * xi`` (UqModel::ScaleParameters, :220-244).
"""
from __future__ import annotations

import os
from typing import Dict, List, Optional, Tuple

import numpy as np

from .model import ModelData


class UqModel:
    """Mirror of nimble::UqModel (src/uq/nimble_uq.h:59-190)."""

    def __init__(self):
        self.samples_file_name = ""
        self.samples_text: Optional[str] = None
        self.nominal_parameters: Dict[int, Dict[str, float]] = {}
        self.parameter_uncertainties: Dict[int, Dict[str, float]] = {}
        self.parameter_order: List[Tuple[int, str]] = []
        self.parameter_samples: List[List[float]] = []
        self.interpolation_coefficients: List[List[float]] = []
        self.parameters: Dict[int, List[Dict[str, float]]] = {}
        self.nexact_samples = 0
        self.napprox_samples = 0
        self.initialized = False

    # -- deck input
    def ParseConfiguration(self, line: str) -> None:
        """``uq model: file <name>`` (src/uq/nimble_uq.cc:92-107)."""
        items = line.split()
        if len(items) < 2 or items[0] != "file":
            raise ValueError(f"uq model: expected 'file <samples file>', got '{line}'")
        self.samples_file_name = items[1]

    def ParseBlockInput(self, line: str, block_id: int, nominal_parameters: Dict[str, float]) -> None:
        """``uq parameters: <material> <name> <range> ...`` for one block
        (src/uq/nimble_uq.cc:110-143): uncertain parameters in line order."""
        if line == "none":
            raise ValueError(f"no uq parameters are specified for the material of block {block_id}")
        items = line.split()
        if len(items) % 2:
            raise ValueError(f"uq parameters of block {block_id}: expected <name> <range> pairs, got '{line}'")
        for i in range(0, len(items), 2):
            self.parameter_uncertainties.setdefault(block_id, {})[items[i]] = float(items[i + 1])
            self.parameter_order.append((block_id, items[i]))
        self.nominal_parameters[block_id] = dict(nominal_parameters)

    # -- samples
    def ReadSamples(self, text: Optional[str] = None) -> None:
        """Parse the samples file (src/uq/nimble_uq.cc:145-218), with the
        reference's checks and messages."""
        name = self.samples_file_name
        if text is None:
            with open(name) as f:
                text = f.read()
        lines = text.splitlines()
        npar = self.GetNumParameters()

        def line_at(k: int) -> str:
            if k >= len(lines):
                raise ValueError(f"UQ file {name} ends before line {k + 1}")
            return lines[k]

        first = line_at(0).split()
        if len(first) < 2:
            raise ValueError(f"UQ file {name}: line 1 must hold the numbers of exact and approximate samples")
        self.nexact_samples, self.napprox_samples = int(first[0]), int(first[1])
        self.parameter_samples, self.interpolation_coefficients = [], []
        ln = 1
        for _ in range(self.nexact_samples):
            items = line_at(ln).split()
            ln += 1
            if len(items) != npar:
                raise ValueError(f"UQ file {name} has not enough exact sample parameters on line {ln} "
                                 f"expecting {npar} found {len(items)}")
            self.parameter_samples.append([float(x) for x in items])
        ncoeff = self.nexact_samples + 1
        for _ in range(self.napprox_samples):
            line = line_at(ln)
            ln += 1
            if ":" not in line:
                raise ValueError(f"UQ file {name} has wrong format (no colon) on line {ln}")
            params, coeffs = line.split(":", 1)
            items = params.split()
            if len(items) != npar:
                raise ValueError(f"UQ file {name} has {len(items)} parameters (preceding :) on line {ln} "
                                 f"expecting {npar}")
            self.parameter_samples.append([float(x) for x in items])
            items = coeffs.split()
            if len(items) != ncoeff:
                raise ValueError(f"UQ file {name} has {len(items)} coefficients on line {ln} expecting {ncoeff}")
            self.interpolation_coefficients.append([float(x) for x in items])

    def ScaleParameters(self) -> None:
        """p = p0 + range * xi per sample (src/uq/nimble_uq.cc:220-244)."""
        ns = self.GetNumSamples()
        self.parameters = {b: [dict(p0) for _ in range(ns)] for b, p0 in self.nominal_parameters.items()}
        for i, (b, name) in enumerate(self.parameter_order):
            dpdxi = self.parameter_uncertainties[b][name]
            for s in range(ns):
                p0 = self.parameters[b][s][name]
                self.parameters[b][s][name] = p0 + dpdxi * self.parameter_samples[s][i]

    def WriteSamples(self, path: str = "parameter_samples.dat") -> None:
        """Scaled parameters, one sample per line (src/uq/nimble_uq.cc:246-268)."""
        with open(path, "w") as f:
            f.write("# " + "".join(f"{name}_{b} " for b, name in self.parameter_order) + "\n")
            for s in range(self.GetNumSamples()):
                f.write("".join("%8g " % self.parameters[b][s][name] for b, name in self.parameter_order) + "\n")

    # -- device ensemble
    def Initialize(self, model: ModelData, samples_text: Optional[str] = None,
                   write_samples: Optional[str] = None) -> None:
        """ReadSamples, ScaleParameters, (WriteSamples), then the device
        allocation of every sample trajectory (src/uq/nimble_uq.cc:270-300).
        Every block needs uq parameters: the reference indexes each block's
        sample parameters when it evaluates the exact samples
        (src/uq/nimble_uq_model_data.cc:180-186)."""
        self.ReadSamples(samples_text)
        self.ScaleParameters()
        if write_samples:
            self.WriteSamples(write_samples)
        ids = [b.block_id for b in model.blocks]
        missing = [b for b in ids if b not in self.parameters]
        if missing:
            raise ValueError(f"uq parameters are missing for block(s) {missing}")
        ne = self.nexact_samples
        K = np.array([[self.parameters[b][s]["bulk_modulus"] for b in ids] for s in range(ne)]).reshape(ne, len(ids))
        G = np.array([[self.parameters[b][s]["shear_modulus"] for b in ids] for s in range(ne)]).reshape(ne, len(ids))
        c = np.array(self.interpolation_coefficients, np.float64).reshape(self.napprox_samples, ne + 1)
        model.SetUqSamples(K, G, c)
        self.initialized = True

    # -- accessors (src/uq/nimble_uq.h:113-155)
    def GetNumSamples(self) -> int:
        return self.nexact_samples + self.napprox_samples

    def GetNumParameters(self) -> int:
        return len(self.parameter_order)

    def GetNumExactSamples(self) -> int:
        return self.nexact_samples

    def Parameters(self, block_id: int, sample: int) -> Dict[str, float]:
        return self.parameters[block_id][sample]


def block_id_of_material(element_blocks, material_key: str) -> int:
    """First block (ascending id) whose material is material_key, else -1
    (Parser::GetBlockIdFromMaterial, src/nimble_parser.h:227-243)."""
    for bid, mkey in sorted(element_blocks):
        if mkey == material_key:
            return bid
    return -1


def uq_from_deck(deck, model: ModelData, write_samples: bool = False) -> Optional[UqModel]:
    """Attach the deck's ensemble to the model (ModelData::InitializeBlocks of
    the UQ build, src/uq/nimble_uq_model_data.cc:64-105): uq parameter lines
    in material-name order (the parser's std::map), the samples file relative
    to the deck's directory.  None when the deck has no UQ keys
    (Parser::HasUq, src/nimble_parser.h:384-391)."""
    if not deck.uq_model or not deck.uq_parameters:
        return None
    uq = UqModel()
    uq.ParseConfiguration(deck.uq_model)
    for mkey, line in sorted(deck.uq_parameters.items()):
        bid = block_id_of_material(deck.element_blocks, mkey)
        if bid < 0:
            raise ValueError(f"uq parameters: material '{mkey}' is not used by any element block")
        uq.ParseBlockInput(line, bid, deck.materials[mkey])
    path = uq.samples_file_name
    if not os.path.isabs(path):
        uq.samples_file_name = os.path.join(deck.directory, path)
    uq.Initialize(model, write_samples=os.path.join(deck.directory, "parameter_samples.dat") if write_samples
                  else None)
    return uq
