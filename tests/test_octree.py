"""Octree indexing: Morton keys, lists and masks of the host octree, bit-exact
against the pure-Python restatement (oracle/tree_ref.py) and the reference's
own tree expectations (tests/tree/tree.cpp:51-83)."""
import numpy as np
import pytest

from iblgf_amd import FLAG_CORRECTION, FLAG_LEAF, FLAG_PHYSICAL, Tree, lib, setups
from oracle import tree_ref as R

RNG = np.random.default_rng(12345)


def test_key_layout_matches_restatement():
    L = lib()
    for _ in range(3000):
        lvl = int(RNG.integers(0, 20))
        x, y, z = (int(v) for v in RNG.integers(-5, (1 << lvl) + 5, 3))
        k = L.iblgf_key_make(x, y, z, lvl)
        assert k == R.make(x, y, z, lvl)
        assert k & 1 == 1 and R.level(k) == lvl
        c = (np.zeros(3, np.int32))
        assert L.iblgf_key_coord(k, c.ctypes.data_as(__import__("ctypes").POINTER(__import__("ctypes").c_int))) == lvl
        m = (1 << lvl) - 1 if lvl else 0
        # coordinates wrap modulo 2^level
        assert tuple(int(v) for v in c) == (x & m, y & m, z & m)
        assert L.iblgf_key_parent(k) == R.parent(k)
        if lvl < 19:
            for i in range(8):
                ck = L.iblgf_key_child(k, i)
                assert ck == R.child(k, i)
                assert R.parent(ck) == k and R.child_number(ck, lvl + 1) == i


def test_key_closed_forms():
    # level-L keys use the top 3L coordinate bits; x is the lowest of each triple
    assert R.make(1, 0, 0, 1) == (1 << 61) | (1 << 2) | 1
    assert R.make(0, 1, 0, 1) == (1 << 62) | (1 << 2) | 1
    assert R.make(0, 0, 1, 1) == (1 << 63) | (1 << 2) | 1
    assert R.make(0, 0, 0, 0) == 1
    # child ids: bit0 = x, bit1 = y, bit2 = z
    k = R.make(3, 5, 6, 4)
    for i in range(8):
        c = R.coord(R.child(k, i))
        assert c == (6 + (i & 1), 10 + ((i >> 1) & 1), 12 + ((i >> 2) & 1))
    # wrap-around: -1 aliases 2^L - 1
    assert R.make(-1, 0, 0, 3) == R.make(7, 0, 0, 3)
    assert R.clog2(19) == 5 and R.clog2(4) == 3


def test_influence_keys_match():
    L = lib()
    import ctypes

    buf = (ctypes.c_uint64 * 216)()
    for _ in range(200):
        lvl = int(RNG.integers(1, 10))
        x, y, z = (int(v) for v in RNG.integers(0, 1 << lvl, 3))
        k = R.make(x, y, z, lvl)
        n = L.iblgf_key_influence(k, buf)
        ref = R.influence_keys(k)
        assert list(buf[:n]) == ref
        assert len(ref) <= 216
        for ik in ref:
            c, ci = R.coord(k), R.coord(ik)
            assert max(abs(ci[d] - c[d]) for d in range(3)) > 1


def test_reference_tree_expectations():
    # tests/tree/tree.cpp:51-83: a 5^3-block tree has base level 3, depth 4,
    # every base block is a leaf and no coarser octant is.
    t = Tree((0, 0, 0), (40, 40, 40), 8, 1.0, [(0, 0, 0, 40, 40, 40)]).init_refine(0).build_masks()
    info = t.info()
    assert info["base_level"] == 3 and info["depth"] == 4
    a = t.export()
    lvl = a["level"]
    assert np.all((a["flags"][lvl == 3] & FLAG_LEAF) != 0)
    assert np.all((a["flags"][lvl < 3] & FLAG_LEAF) == 0)
    assert (lvl == 3).sum() == 125


def _trees():
    g = setups.UniformGaussians()
    v = setups.VortexRings(seed=7, block_extent=8, base_blocks=4, nref=2, nrings=3, centre_range=4.0,
                           radius=(4.0, 8.0), core=(1.5, 2.5), refinement_factor=0.5)
    w = setups.VortexRings(seed=7, block_extent=6, base_blocks=6, nref=2, nrings=4, centre_range=6.0,
                           radius=(4.0, 8.0), core=(1.5, 2.5), refinement_factor=0.5)
    return {"uniform": g, "rings": v, "rings_n8": w}


@pytest.fixture(scope="module")
def trees():
    return _trees()


@pytest.mark.parametrize("name", ["uniform", "rings", "rings_n8"])
def test_lists_bit_exact(trees, name):
    t = trees[name]
    a = t.arrays
    nbr, infl = R.lists(a["key"])
    assert np.array_equal(a["neighbor"], np.array(nbr, np.int32))
    for s in range(len(a["key"])):
        k = int(a["influence_num"][s])
        assert k <= 189
        assert list(a["influence"][s][:k]) == infl[s]
        assert np.all(a["influence"][s][k:] == -1)


@pytest.mark.parametrize("name", ["uniform", "rings", "rings_n8"])
def test_masks_bit_exact(trees, name):
    t = trees[name]
    a = t.arrays
    S, T = R.amr2amr_masks(a, t.info["base_level"], t.info["depth"])
    hi = ~((1 << 15) - 1) & ((1 << 64) - 1)  # AMR2AMR indices 15..
    for s in range(len(a["key"])):
        assert int(a["mask_src"][s]) & hi == S[s] & hi
        assert int(a["mask_tgt"][s]) & hi == T[s] & hi


@pytest.mark.parametrize("name", ["uniform", "rings", "rings_n8"])
def test_tree_invariants(trees, name):
    t = trees[name]
    a = t.arrays
    nb = len(a["key"])
    keys = [int(k) for k in a["key"]]
    # slots are level-major and key-sorted within a level
    order = sorted(range(nb), key=lambda s: (int(a["level"][s]), keys[s]))
    assert order == list(range(nb))
    for s in range(nb):
        assert a["neighbor"][s][13] == s
        for i, nb_s in enumerate(a["neighbor"][s]):
            if nb_s < 0:
                continue
            # symmetry: s is the mirrored neighbour of its neighbour (no wrap in these trees)
            assert a["neighbor"][nb_s][26 - i] == s
        for c in a["children"][s]:
            if c >= 0:
                assert a["parent"][c] == s and a["level"][c] == a["level"][s] + 1
    # 2:1 balance and correction buffer: every physical block above the base
    # level has all 26 neighbours, and leaves only touch leaves one level apart
    base = t.info["base_level"]
    for s in range(nb):
        if a["level"][s] > base and (a["flags"][s] & FLAG_PHYSICAL):
            assert np.all(a["neighbor"][s] >= 0)
    leaves = np.nonzero(a["flags"] & FLAG_LEAF)[0]
    assert len(leaves) == t.info["nleaves"]
    assert np.any(a["flags"] & FLAG_CORRECTION) or name == "uniform"


def test_adaptive_vortex_ring_tree_size():
    v = setups.VortexRings()
    assert v.info["nblocks"] == 4290 and v.info["nleaves"] == 1933
    assert v.info["base_level"] == 5 and v.info["depth"] == 9


@pytest.mark.parametrize("name", ["uniform", "rings", "rings_n8"])
def test_tree_import_round_trip(trees, name):
    """A tree handed over by keys and flags rebuilds the same lists and masks."""
    t = trees[name]
    a = t.arrays
    # base shift = bounding-box base / block extent = root block base / e (root coord is 0)
    assert int(a["level"][0]) == 0
    shift = [int(a["bbase"][0][d] // t.block_extent) for d in range(3)]
    perm = np.random.default_rng(5).permutation(len(a["key"]))
    u = Tree.from_keys(shift, t.block_extent, t.info["base_level"], t.info["dx_base"], a["key"][perm],
                       a["flags"][perm] & 3)
    u.build_masks()
    b = u.export()
    for k in ("key", "level", "coord", "bbase", "parent", "children", "neighbor", "influence", "influence_num",
              "mask_src", "mask_tgt"):
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(a["flags"] & 3, b["flags"] & 3)
