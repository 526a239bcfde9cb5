"""Extract golden fixtures from NimbleSM's own regression data.

Reads (as data only) the genesis meshes, input decks and ``*.gold.e`` Exodus
results under ``/root/reference/test/dynamics`` and ``/root/reference/test/uq``
with ``scipy.io.netcdf_file`` and writes one compressed ``.npz`` per deck next
to this script.  Nothing of
the reference is imported or executed.  Run in a container that has the
reference tree:

    python tests/golden/make_golden.py

Fixture contents (all 0-based):
  X [N,3], blk_ids [B], blk_off [B+1], conn [E,8] (blocks in ascending id),
  rho/K/G [B], bc_node/bc_coord/bc_kind/bc_value (deck order x node-set order),
  t_final, num_load_steps, output_frequency,
  gold_time [T], gold_<nodal var> [T,N], gold_<elem var> [T,E] (element order
  as in conn).
UQ decks (uq_wave; its gold file is not in the reference checkout, so no
gold_* arrays) add: uq_model (the 'uq model' value), uq_material / uq_line
(the 'uq parameters' lines: material key, rest of the line), uq_samples_text
(the samples file named by uq_model, as text), output_fields.
"""
from __future__ import annotations

import os
import re
import sys

import numpy as np
from scipy.io import netcdf_file

REF = "/root/reference/test"
OUT = os.path.dirname(os.path.abspath(__file__))

DECKS = {
    "single_elem_native_neohookean": ("dynamics/single_elem_complex_displacement", "single_elem.g"),
    "simple_deformation_modes": ("dynamics/simple_deformation_modes", "simple_deformation_modes.g"),
    "rigid_body_motion": ("dynamics/rigid_body_motion", "rigid_body_motion.g"),
    "wave_in_bar": ("dynamics/wave_in_bar", "wave_in_bar.g"),
    "brick_with_fibers": ("dynamics/brick_with_fibers", "brick_with_fibers.g"),
    "notched_plate_native_neohookean": ("dynamics/notched_plate_native_neohookean",
                                        "notched_plate_native_neohookean.g"),
    "uq_wave": ("uq/uq_wave", "uq_wave.g"),
}

COORD = {"x": 0, "y": 1, "z": 2}
KIND = {"initial_velocity": 0, "prescribed_velocity": 1, "prescribed_displacement": 2}


def _names(f, var):
    if var not in f.variables:
        return []
    return [b"".join(r).decode("latin1").strip("\x00").strip() for r in f.variables[var][:]]


def read_genesis(path):
    f = netcdf_file(path, "r", mmap=False)
    X = np.stack([f.variables[c][:].astype(np.float64) for c in ("coordx", "coordy", "coordz")], axis=1)
    blk_ids = [int(v) for v in f.variables["eb_prop1"][:]]
    blocks = {}
    for i, bid in enumerate(blk_ids):
        blocks[bid] = (i + 1, f.variables[f"connect{i + 1}"][:].astype(np.int64) - 1)
    ns = {}
    if "ns_prop1" in f.variables:
        for i, nid in enumerate(f.variables["ns_prop1"][:]):
            ns[int(nid)] = f.variables[f"node_ns{i + 1}"][:].astype(np.int64) - 1
    f.close()
    return X, blocks, ns


def parse_deck(path):
    deck = {"materials": {}, "blocks": [], "bcs": [], "uq": {}}
    for line in open(path):
        line = line.split("#")[0].strip()
        if ":" not in line:
            continue
        key, val = [s.strip() for s in line.split(":", 1)]
        if key == "final time":
            deck["t_final"] = float(val)
        elif key == "initial time":
            deck["t_initial"] = float(val)
        elif key == "number of load steps":
            deck["num_load_steps"] = int(val)
        elif key == "output frequency":
            deck["output_frequency"] = int(val)
        elif key == "material parameters":
            tok = val.split()
            name, model = tok[0], tok[1]
            assert model == "neohookean", model
            p = {tok[i]: float(tok[i + 1]) for i in range(2, len(tok), 2)}
            deck["materials"][name] = p
        elif key == "element block":
            b, m = val.split()
            deck["blocks"].append((int(re.sub(r"\D", "", b)), m))
        elif key == "boundary condition":
            kind, ns, coord, mag = val.split()
            deck["bcs"].append((KIND[kind], int(re.sub(r"\D", "", ns)), COORD[coord], float(mag)))
        elif key == "uq parameters":
            mat, rest = val.split(None, 1)
            deck["uq"][mat] = rest.strip()
        elif key == "uq model":
            deck["uq_model"] = val
        elif key == "output fields":
            deck["output_fields"] = val
    deck.setdefault("t_initial", 0.0)
    deck.setdefault("output_frequency", 1)
    return deck


def make(name, subdir, mesh):
    d = os.path.join(REF, subdir)
    X, blocks, ns = read_genesis(os.path.join(d, mesh))
    deck = parse_deck(os.path.join(d, name + ".in"))
    gold_path = os.path.join(d, name + ".gold.e")
    gold = netcdf_file(gold_path, "r", mmap=False) if os.path.exists(gold_path) else None

    mat_of_block = dict(deck["blocks"])
    order = sorted(blocks)  # ascending block id, as std::map in ModelData
    conn, blk_off, rho, K, G = [], [0], [], [], []
    for bid in order:
        conn.append(blocks[bid][1])
        blk_off.append(blk_off[-1] + len(blocks[bid][1]))
        p = deck["materials"][mat_of_block[bid]]
        rho.append(p["density"])
        K.append(p["bulk_modulus"])
        G.append(p["shear_modulus"])
    conn = np.concatenate(conn).astype(np.int32)

    bn, bc, bk, bv = [], [], [], []
    for kind, nsid, coord, mag in deck["bcs"]:
        for n in ns[nsid]:
            bn.append(n)
            bc.append(coord)
            bk.append(kind)
            bv.append(mag)

    out = dict(
        X=X, blk_ids=np.array(order), blk_off=np.array(blk_off, np.int64), conn=conn,
        rho=np.array(rho), K=np.array(K), G=np.array(G),
        bc_node=np.array(bn, np.int32), bc_coord=np.array(bc, np.int32), bc_kind=np.array(bk, np.int32),
        bc_value=np.array(bv), t_final=deck["t_final"], t_initial=deck["t_initial"],
        num_load_steps=deck["num_load_steps"], output_frequency=deck["output_frequency"],
    )
    if deck["uq"]:
        mats = sorted(deck["uq"])
        out.update(uq_model=deck["uq_model"], uq_material=np.array(mats), uq_line=np.array([deck["uq"][m] for m in mats]),
                   output_fields=deck.get("output_fields", ""))
        fname = deck["uq_model"].split()[1]
        out["uq_samples_text"] = open(os.path.join(d, fname)).read()
    if gold is None:
        path = os.path.join(OUT, name + ".npz")
        np.savez_compressed(path, **out)
        print(f"{name}: N={X.shape[0]} E={conn.shape[0]} (no gold file) -> {path}")
        return
    out["gold_time"] = gold.variables["time_whole"][:].astype(np.float64)
    # gold results: nodal vars by name, element vars concatenated over blocks in ascending id
    gX = np.stack([gold.variables[c][:] for c in ("coordx", "coordy", "coordz")], axis=1)
    assert np.array_equal(gX, X), "gold mesh differs from genesis mesh"
    for i, nm in enumerate(_names(gold, "name_nod_var")):
        out["gold_" + nm] = gold.variables[f"vals_nod_var{i + 1}"][:].astype(np.float64)
    gblk = [int(v) for v in gold.variables["eb_prop1"][:]]
    for i, nm in enumerate(_names(gold, "name_elem_var")):
        parts = []
        for bid in order:
            bi = gblk.index(bid) + 1
            key = f"vals_elem_var{i + 1}eb{bi}"
            if key not in gold.variables:
                parts = None
                break
            parts.append(gold.variables[key][:].astype(np.float64))
        if parts is not None:
            out["gold_" + nm] = np.concatenate(parts, axis=1)
    gold.close()
    path = os.path.join(OUT, name + ".npz")
    np.savez_compressed(path, **out)
    print(f"{name}: N={X.shape[0]} E={conn.shape[0]} vars={sum(k.startswith('gold_') for k in out)} -> {path}")


def main(argv):
    names = argv[1:] or list(DECKS)
    for n in names:
        make(n, *DECKS[n])


if __name__ == "__main__":
    main(sys.argv)
