"""Adaptation of the host tree (iblgf_tree_adapt: Decomposition::adapt_decoposition,
decomposition.hpp:182-470, on one rank) and the oracle's field remap
(Ifherk::adapt, ifherk.hpp:1018-1076).

The tree rules are pinned structurally: a second adaptation without level
changes is the identity; refined leaves get eight children; the 2:1 balance
holds (every block's same-level neighbours have parents); every physical block
above the base level has all 26 same-level neighbours (the correction layer,
mark_correction, domain.hpp:617-735); a sibling group asked to coarsen
leaves its parent a leaf.  The remap is pinned by geometry: blocks kept by key
are unchanged and the children of a refined block hold the interpolation of
a quadratic face field at their own positions.
"""
import ctypes

import numpy as np
import pytest

from iblgf_amd import FLAG_CORRECTION, FLAG_LEAF, Tree, lib, setups

PHYSICAL = 8


def _coord(k):
    c = (ctypes.c_int * 3)()
    lib().iblgf_key_coord(ctypes.c_uint64(int(k)), c)
    return list(c)


def _make(key=None):
    """Base region of 8^3 blocks in a 16^3-block box, refined twice around the centre."""
    t = Tree((-64,) * 3, (128,) * 3, 8, 1.0 / 64, [(-32, -32, -32, 64, 64, 64)])

    def cond(b, ext, rl, diff):
        c = [(b[d] + ext / 2) / 2 ** rl for d in range(3)]
        return np.linalg.norm(c) < 14 / (1 + rl)

    t.init_refine(2, cond)
    t.build_masks()
    return t


def _state(t):
    a = t.export()
    return {int(k): (int(f) & 3, int(l)) for k, f, l in zip(a["key"], a["flags"], a["level"])}, a


def _check_tree(a, info):
    keys = {int(k): s for s, k in enumerate(a["key"])}
    bl = info["base_level"]
    for s in range(info["nblocks"]):
        l = int(a["level"][s])
        if l <= bl:
            continue
        k = int(a["key"][s])
        c = _coord(k)
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    nk = int(lib().iblgf_key_make(c[0] + dx, c[1] + dy, c[2] + dz, l))
                    pk = int(lib().iblgf_key_parent(ctypes.c_uint64(nk)))
                    assert pk in keys, ("2:1 balance", k, (dx, dy, dz))
                    if a["flags"][s] & PHYSICAL:
                        assert nk in keys, ("correction layer", k, (dx, dy, dz))


def _settle(t):
    """Adapt without level changes until no base-level block is added: each
    pass adds the base blocks within two blocks of a non-correction one
    (baseBlockBufferNumber = 2), and the new outer ring becomes correction, so
    the base mesh grows by a ring per pass up to the bounding box."""
    for i in range(12):
        upd = t.adapt({})
        t.build_masks()
        if not upd:
            return i
    raise AssertionError("the base mesh kept growing")


def test_adapt_without_changes_reaches_a_fixed_point():
    t = _make()
    passes = _settle(t)
    assert passes >= 1
    s1, a1 = _state(t)
    assert t.adapt({}) is False
    t.build_masks()
    s2, a2 = _state(t)
    assert s1 == s2
    info = t.info()
    _check_tree(a2, info)
    # the base mesh fills the bounding box (16 blocks a side)
    assert np.sum(a2["level"] == info["base_level"]) == 16 ** 3


def test_refinement_and_coarsening_keep_the_tree_rules():
    t = _make()
    _settle(t)
    s, a = _state(t)
    info = t.info()
    bl = info["base_level"]
    fin = [int(k) for k, f, l in zip(a["key"], a["flags"], a["level"])
           if (f & FLAG_LEAF) and not (f & FLAG_CORRECTION) and l == bl + 2]
    # refine the leaves nearest the centre
    fin.sort(key=lambda k: np.linalg.norm(np.array(_coord(k)) - 31.5))
    ref = fin[:3]
    t.adapt({k: 1 for k in ref})
    t.build_masks()
    s2, a2 = _state(t)
    info2 = t.info()
    assert info2["depth"] == info["depth"] + 1
    keys2 = {int(k): i for i, k in enumerate(a2["key"])}
    for k in ref:
        assert s2[k][0] & FLAG_LEAF == 0
        kids = [int(c) for c in a2["children"][keys2[k]] if c >= 0]
        assert len(kids) == 8 and all(s2[int(a2["key"][c])][0] & FLAG_LEAF for c in kids)
    _check_tree(a2, info2)
    # coarsen: every finest sibling group of non-correction leaves away from the refined spot
    groups = {}
    for i, (k, f, l) in enumerate(zip(a2["key"], a2["flags"], a2["level"])):
        if l == bl + 2 and (f & FLAG_LEAF) and not (f & FLAG_CORRECTION):
            groups.setdefault(int(a2["parent"][i]), []).append(int(k))
    full = [p for p, ks in groups.items() if len(ks) == 8]
    assert full
    far = max(full, key=lambda p: np.linalg.norm(np.array(_coord(int(a2["key"][p]))) - 15.75))
    t.adapt({k: -1 for k in groups[far]})
    t.build_masks()
    s3, a3 = _state(t)
    pk = int(a2["key"][far])
    assert s3[pk][0] & FLAG_LEAF
    assert not any(k in s3 and (s3[k][0] & FLAG_LEAF) and not (s3[k][0] & FLAG_CORRECTION) for k in groups[far])
    _check_tree(a3, t.info())


def test_remap_keeps_blocks_and_interpolates_new_ones(oracle_mod):
    t = _make()
    _settle(t)
    _, a = _state(t)
    info, n = t.info(), t.n
    bl = info["base_level"]
    fin = [int(k) for k, f, l in zip(a["key"], a["flags"], a["level"])
           if (f & FLAG_LEAF) and not (f & FLAG_CORRECTION) and l == bl + 2]
    fin.sort(key=lambda k: np.linalg.norm(np.array(_coord(k)) - 31.5))
    poly = lambda X, Y, Z: 0.3 + 0.7 * X - 0.4 * Y + 0.9 * Z + 0.05 * X * Y - 0.03 * Z * Z

    def sample(arrays, inf):
        out = []
        for c in range(3):
            f = np.zeros((inf["nblocks"], n, n, n))
            for s in range(inf["nblocks"]):
                P = setups.mesh_points(arrays, s, n, inf["dx_base"], inf["base_level"], 1, c) / inf["dx_base"]
                f[s] = poly(P[..., 0], P[..., 1], P[..., 2])
            out.append(f)
        return out

    u = sample(a, info)
    t.adapt({k: 1 for k in fin[:3]})
    t.build_masks()
    a2 = t.export()
    info2 = t.info()
    ot = oracle_mod.Tree(a2, info2, n)
    got, parents = oracle_mod.remap(a, ot, u, [1, 1, 1], [0, 1, 2])
    assert len(parents) == 3
    old = {int(k): s for s, k in enumerate(a["key"])}
    exact = sample(a2, info2)
    inner = (slice(1, -1),) * 3
    new_blocks = 0
    for s in range(info2["nblocks"]):
        k = int(a2["key"][s])
        if k in old and int(a2["parent"][s]) not in parents:
            # kept blocks: values kept, except the halos that the fills refresh
            for c in range(3):
                assert np.array_equal(got[c][s][inner], u[c][old[k]][inner])
        elif k not in old:
            new_blocks += 1
            if int(a2["parent"][s]) >= 0 and int(a2["key"][int(a2["parent"][s])]) in fin[:3]:
                for c in range(3):
                    assert np.max(np.abs(got[c][s][inner] - exact[c][s][inner])) < 1e-9
    assert new_blocks >= 24
