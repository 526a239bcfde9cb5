"""Host tile builder of the tile-local mixed-precision assembly
(nimblesm_amd/csrc/nsm_tiles.cpp), compiled with g++ from its source and
checked on the CPU: every element lands in exactly one tile of its block,
the local connectivity maps back to the mesh under the builder's node
renumbering, colours separate elements that share a node, tiles flagged for
the per-corner gather hold each node at most once per corner index, and
interior nodes + ordered boundary partials reproduce the element-force
scatter sum of every node exactly once."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MAX_COLORS = 32


@pytest.fixture(scope="module")
def tiles_lib(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("tiles") / "libtiles.so")
    subprocess.run(
        ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-o", out,
         os.path.join(ROOT, "tests", "tiles_shim.cpp"),
         os.path.join(ROOT, "nimblesm_amd", "csrc", "nsm_tiles.cpp")],
        check=True,
    )
    lib = C.CDLL(out)
    lib.tiles_build.restype = C.c_void_p
    lib.tiles_build.argtypes = [C.c_int64, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int]
    lib.tiles_sizes.argtypes = [C.c_void_p, C.c_void_p]
    lib.tiles_copy.argtypes = [C.c_void_p] + [C.c_void_p] * 15
    lib.tiles_free.argtypes = [C.c_void_p]
    return lib


def build(lib, X, conn, blk_off, tile_elems=512):
    X = np.ascontiguousarray(X, np.float64)
    conn = np.ascontiguousarray(conn, np.int32)
    off = np.ascontiguousarray(blk_off, np.int64)
    h = lib.tiles_build(len(X), X.ctypes.data, conn.ctypes.data, len(off) - 1, off.ctypes.data, tile_elems)
    assert h, "tile build failed"
    sz = np.zeros(5, np.int64)
    lib.tiles_sizes(h, sz.ctypes.data)
    T, S, B, BE, _ = (int(x) for x in sz)
    E = len(conn)
    t = dict(
        perm=np.zeros(E, np.int64), elem_off=np.zeros(T + 1, np.int32), node_off=np.zeros(T + 1, np.int32),
        nodes=np.zeros(S, np.int32), node_info=np.zeros(S, np.uint8), lconn=np.zeros((E, 8), np.uint16),
        color=np.zeros(E, np.uint8), num_colors=np.zeros(T, np.uint8), bnd_nodes=np.zeros(B, np.int32),
        bnd_ptr=np.zeros(B + 1, np.int32), bnd_ent=np.zeros(BE, np.int32), slot_dst=np.zeros(S, np.int32),
        node_new=np.zeros(len(X), np.int32), int_off=np.zeros(T + 1, np.int32), gather=np.zeros(T, np.uint8),
    )
    lib.tiles_copy(h, *(t[k].ctypes.data for k in ("perm", "elem_off", "node_off", "nodes", "node_info", "lconn",
                                                   "color", "num_colors", "bnd_nodes", "bnd_ptr", "bnd_ent",
                                                   "slot_dst", "node_new", "int_off", "gather")))
    lib.tiles_free(h)
    t["num_interior"] = int(sz[4])
    t["tile_elems"] = tile_elems
    return t


def check_invariants(t, N, conn_mesh, blk_off):
    E = len(conn_mesh)
    perm, eo, no = t["perm"], t["elem_off"], t["node_off"]
    T = len(eo) - 1
    assert np.array_equal(np.sort(perm), np.arange(E))
    # node renumbering: a permutation; tile t's interior nodes are the
    # contiguous new ids int_off[t]..int_off[t+1] in its first slots, boundary
    # nodes follow all interior ones
    nn = t["node_new"]
    assert np.array_equal(np.sort(nn), np.arange(N))
    conn = nn[conn_mesh]
    io = t["int_off"]
    for ti in range(T):
        k = io[ti + 1] - io[ti]
        assert np.array_equal(t["nodes"][no[ti]:no[ti] + k], np.arange(io[ti], io[ti + 1]))
        assert np.all(t["node_info"][no[ti]:no[ti] + k] & 1) and not np.any(t["node_info"][no[ti] + k:no[ti + 1]] & 1)
    assert np.array_equal(t["bnd_nodes"], np.arange(io[-1], N))
    # each block's elements stay in the block's range; tiles never straddle blocks
    for b in range(len(blk_off) - 1):
        seg = perm[blk_off[b]:blk_off[b + 1]]
        assert seg.size == 0 or (seg.min() >= blk_off[b] and seg.max() < blk_off[b + 1])
    assert eo[0] == 0 and eo[-1] == E and np.all(np.diff(eo) > 0) and np.all(np.diff(eo) <= t["tile_elems"])
    assert set(blk_off.tolist()) <= set(eo.tolist())
    interior_of = {}
    for ti in range(T):
        tn = t["nodes"][no[ti]:no[ti + 1]]
        assert 0 < len(tn) <= 2 * t["tile_elems"] and np.all(np.diff(tn) > 0)
        el = np.arange(eo[ti], eo[ti + 1])
        lc = t["lconn"][el].astype(np.int64)
        assert np.array_equal(tn[lc], conn[perm[el]])  # local connectivity maps back to the mesh
        assert np.array_equal(np.unique(conn[perm[el]]), tn)  # slots = the tile's distinct nodes
        col = t["color"][el] & 0x1F
        assert np.all(np.diff(col.astype(int)) >= 0)  # elements of a tile in colour order
        repeated = np.array([len(np.unique(r)) < 8 for r in conn[perm[el]]])
        assert np.array_equal((t["color"][el] & 0x80) != 0, repeated)  # flag of elements listing a node twice
        assert col.max() < t["num_colors"][ti] <= MAX_COLORS
        for c in range(int(t["num_colors"][ti])):
            members = lc[col == c]
            flat = np.concatenate([np.unique(r) for r in members]) if len(members) else np.zeros(0)
            assert len(flat) == len(np.unique(flat)), "two elements of one colour share a node"
        # gather flag <=> no slot occurs twice at one corner index
        unique_corners = all(len(np.unique(lc[:, j])) == len(lc) for j in range(8))
        assert bool(t["gather"][ti]) == unique_corners
        for s in range(no[ti], no[ti + 1]):
            if t["node_info"][s] & 1:
                assert int(t["nodes"][s]) not in interior_of
                interior_of[int(t["nodes"][s])] = ti
    # interior: every incident element is in the tile
    deg = np.bincount(conn.reshape(-1), minlength=N)
    tile_of_pos = np.repeat(np.arange(T), np.diff(eo))
    tile_of_elem = np.empty(E, np.int64)
    tile_of_elem[perm] = tile_of_pos
    for n, ti in interior_of.items():
        inc = np.nonzero((conn == n).any(axis=1))[0]
        assert np.all(tile_of_elem[inc] == ti) and deg[n] > 0
    assert t["num_interior"] == len(interior_of)
    bnd = t["bnd_nodes"]
    assert np.all(np.diff(bnd) > 0)
    assert set(bnd.tolist()).isdisjoint(interior_of) and len(bnd) + len(interior_of) == N
    return interior_of


def tiled_assembly(t, N, felem_internal):
    """Sum element forces the way the kernels do: per tile, in ascending
    corner index for gather tiles and in colour order otherwise; interior
    nodes complete, the others through partials stored at slot_dst (by node,
    ascending tile order).  felem_internal is [E][8][3] in internal (tile)
    element order."""
    f = np.full((N, 3), np.nan)
    eo, no = t["elem_off"], t["node_off"]
    partials = np.full((len(t["bnd_ent"]), 3), np.nan)
    for ti in range(len(eo) - 1):
        acc = np.zeros((no[ti + 1] - no[ti], 3))
        el = np.arange(eo[ti], eo[ti + 1])
        if t["gather"][ti]:
            for j in range(8):
                for e in el:  # at most one element holds a slot at corner j
                    acc[t["lconn"][e, j]] += felem_internal[e, j]
        else:
            for c in range(int(t["num_colors"][ti])):
                for e in el[(t["color"][el] & 0x1F) == c]:
                    for j in range(8):
                        acc[t["lconn"][e, j]] += felem_internal[e, j]
        for ls in range(len(acc)):
            s = no[ti] + ls
            if t["node_info"][s] & 1:
                assert t["slot_dst"][s] == -1
                f[t["nodes"][s]] = acc[ls]
            else:
                partials[t["slot_dst"][s]] = acc[ls]
    for b, n in enumerate(t["bnd_nodes"]):
        p0, p1 = t["bnd_ptr"][b], t["bnd_ptr"][b + 1]
        ent = t["bnd_ent"][p0:p1]
        assert np.all(np.diff(ent) > 0)  # ascending tile order
        assert np.array_equal(t["slot_dst"][ent], np.arange(p0, p1))
        assert np.all(t["nodes"][ent] == n)
        f[n] = partials[p0:p1].sum(axis=0) if p1 > p0 else 0.0
    return f


def cube(n, seed=5):
    X, conn = O.jittered_cube(n, seed, 0.1)
    return X, conn


@pytest.mark.parametrize("tile_elems,nodes_per_tile",
                         [(512, 9 * 9 * 9), (256, 9 * 9 * 5), (128, 9 * 5 * 5), (64, 5 * 5 * 5)])
def test_structured_cube_tiles_are_boxes(tiles_lib, tile_elems, nodes_per_tile):
    n = 16
    X, conn = cube(n)
    blk_off = np.array([0, n ** 3 // 2, n ** 3])
    t = build(tiles_lib, X, conn, blk_off, tile_elems)
    interior = check_invariants(t, len(X), conn, blk_off)
    # Morton tiles of 512 (256, 128, 64) elements on the jittered lattice are
    # 8x8x8 (8x8x4, 8x4x4, 4x4x4) boxes of 9x9x9 (9x9x5, 9x5x5, 5x5x5) nodes
    assert len(t["elem_off"]) - 1 == n ** 3 // tile_elems
    assert np.all(np.diff(t["node_off"]) == nodes_per_tile)
    assert np.all(t["num_colors"] == 8)
    assert np.all(t["gather"] == 1)  # a lattice holds every node once per corner index
    assert len(interior) > 0.3 * len(X)


def test_assembly_covers_every_incidence_once(tiles_lib):
    n = 12
    X, conn = cube(n, seed=9)
    blk_off = np.array([0, 700, n ** 3])
    t = build(tiles_lib, X, conn, blk_off)
    check_invariants(t, len(X), conn, blk_off)
    rng = np.random.default_rng(1)
    felem = rng.normal(size=(len(conn), 8, 3))
    ref = np.zeros((len(X), 3))
    np.add.at(ref, t["node_new"][conn].reshape(-1), felem.reshape(-1, 3))
    got = tiled_assembly(t, len(X), felem[t["perm"]])
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12)


def test_unstructured_permuted_mesh_with_orphans_and_collapsed_elements(tiles_lib):
    n = 10
    X, conn = cube(n, seed=3)
    rng = np.random.default_rng(7)
    # renumber nodes randomly, shuffle elements, add 5 orphan nodes
    N = len(X) + 5
    new_id = rng.permutation(N)[: len(X)]
    Xn = np.zeros((N, 3))
    Xn[new_id] = X
    Xn[np.setdiff1d(np.arange(N), new_id)] = rng.uniform(size=(5, 3))
    conn = new_id[conn][rng.permutation(len(conn))].astype(np.int32)
    # collapse one face of a few elements (repeated node ids within an element)
    for e in (0, 17, 400):
        conn[e, 5] = conn[e, 4]
        conn[e, 6] = conn[e, 7]
    # relabel the corners of some elements (a quarter turn about their z axis):
    # their tiles hold nodes twice at one corner index and fall back to colours
    rot = np.arange(500, 700)
    conn[rot] = conn[rot][:, [1, 2, 3, 0, 5, 6, 7, 4]]
    blk_off = np.array([0, 1, 333, 333, len(conn)])  # a one-element block and an empty block
    t = build(tiles_lib, Xn, conn, blk_off, 128)
    check_invariants(t, N, conn, blk_off)
    assert 0 < int(t["gather"].sum()) < len(t["gather"])  # both assembly orders occur
    felem = rng.normal(size=(len(conn), 8, 3))
    ref = np.zeros((N, 3))
    np.add.at(ref, t["node_new"][conn].reshape(-1), felem.reshape(-1, 3))
    got = tiled_assembly(t, N, felem[t["perm"]])
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12)


def test_high_valence_star_needs_many_colours(tiles_lib):
    # 40 elements sharing node 0: more than 32 colours inside one tile -> refused
    rng = np.random.default_rng(2)
    E = 40
    X = np.vstack([np.zeros((1, 3)), rng.uniform(size=(7 * E, 3))])
    conn = np.hstack([np.zeros((E, 1), np.int32), 1 + np.arange(7 * E, dtype=np.int32).reshape(E, 7)])
    Xc = np.ascontiguousarray(X)
    cc = np.ascontiguousarray(conn)
    off = np.array([0, E], np.int64)
    h = tiles_lib.tiles_build(len(X), Xc.ctypes.data, cc.ctypes.data, 1, off.ctypes.data, 256)
    assert not h
