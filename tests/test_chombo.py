"""Flow-field output and restart files (Chombo-layout HDF5, Tree::write),
host side: the library's iblgf_chombo_write / iblgf_chombo_read and
iblgf_tree_write_restart / iblgf_tree_read_restart against the oracle's
restatement (oracle/chombo_ref.py) and the reference's own file.

Pins of the restatement:
  - the reference's tests/ns_amr_lgf/configs/ref.hdf5 (HDF5 written by the
    HDF5 library), decoded box by box by the independent reader of
    tests/golden/make_ns_amr_lgf_golden.py, equals what read_u puts on a
    tree of the file's cells, and the library writes the same eight
    components back so that the same decoder gives the same dense arrays;
  - linear fields are reproduced exactly on leaves 1 and 2 levels coarser
    than the file, for cell, face, edge and vertex fields (the geometric
    positions fix the file-index shifts and the face/edge axis rules); a
    quadratic x^2 gives the rule's middle-cell average x^2 + (dx_file/2)^2;
  - the writer's boxes are cubes of 8^r whole blocks that tile the listed
    blocks, and the decoded file equals the fields on every listed block;
  - Tree::write is the pre-order depth-first order, i.e. keys sorted by
    (coordinate bits, level), and a read tree equals the written one."""
import math
import os
import struct
import sys

import numpy as np
import pytest

from iblgf_amd import FLAG_CORRECTION, FLAG_LEAF, MESH_CELL, MESH_EDGE, MESH_FACE, MESH_VERTEX, IblgfError, Tree
from oracle import chombo_ref as C

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
from make_ns_amr_lgf_golden import Hdf5Reader  # noqa: E402

REF_HDF5 = "/root/reference/tests/ns_amr_lgf/configs/ref.hdf5"


def _adaptive(block_extent=6):
    t = Tree((-48,) * 3, (96,) * 3, block_extent, 1.0 / 16, [(-24, -24, -24, 48, 48, 48)])

    def cond(b, ext, rl, diff):
        c = [(b[d] + ext / 2) / 2 ** rl for d in range(3)]
        return np.linalg.norm(np.asarray(c) - 3.0) < 14 / (1 + rl)

    t.init_refine(2, cond)
    t.build_masks()
    return t


def _decode(path):
    """Independent dense decode: {level: (components, {(lo, hi): [ncomp, ext...] arrays}), offsets, attrs}."""
    f = Hdf5Reader(path)
    root = f.group(f.root)
    ra = f.attrs(f.root)
    ncomp = ra["num_components"]
    comps = [ra["component_%d" % i] for i in range(ncomp)]
    out = {}
    for lvl in range(ra["num_levels"]):
        g = f.group(root["level_%d" % lvl])
        baddr, bsize, besize = f.contiguous(g["boxes"])
        boxes = [struct.unpack_from("<6i", f.d, baddr + i * besize) for i in range(bsize // besize)]
        daddr, dsize, _ = f.contiguous(g["data:datatype=0"])
        data = np.frombuffer(f.d[daddr:daddr + dsize], dtype="<f8")
        oa, osz, _ = f.contiguous(f.group(g["data_attributes"])["offsets"])
        offsets = np.frombuffer(f.d[oa:oa + osz], dtype="<i4")
        p, blocks = 0, []
        for b in boxes:
            ext = [b[3 + d] - b[d] + 1 for d in range(3)]
            cnt = ext[0] * ext[1] * ext[2]
            arr = data[p:p + ncomp * cnt].reshape(ncomp, ext[2], ext[1], ext[0])
            blocks.append(((b[:3], b[3:]), arr))
            p += ncomp * cnt
        assert p == data.size
        out[lvl] = dict(blocks=blocks, offsets=offsets, attrs=f.attrs(root["level_%d" % lvl]), data=data)
    return comps, out, ra, besize


def _export(t):
    a = t.export()
    a["block_extent"] = t.block_extent
    return a


def test_write_matches_restatement_and_decodes_to_the_fields(tmp_path):
    t = _adaptive()
    a, info, n, E = _export(t), t.info(), t.n, t.block_extent
    bl = info["base_level"]
    rng = np.random.default_rng(3)
    nb = info["nblocks"]
    u = rng.standard_normal((3, nb, n, n, n))
    p = rng.standard_normal((1, nb, n, n, n))
    path = str(tmp_path / "flow.hdf5")
    t.write_chombo(path, [("u", u), ("p", p)])
    comps, lv, ra, besize = _decode(path)
    assert comps == ["u_0", "u_1", "u_2", "p"] and ra["filetype"] == "VanillaAMRFileType"
    assert ra["data_centering"] == 6 and besize == 216
    ref, names = C.encode(a, bl, [("u", u), ("p", p)])
    assert names == comps and sorted(lv) == sorted(ref)
    slots = C.blocks_list(a, bl, True, True)
    for rl, L in lv.items():
        R = ref[rl]
        assert [tuple(bx) for bx, _ in L["blocks"]] == [(tuple(lo), tuple(hi)) for lo, hi in R["boxes"]]
        assert np.array_equal(L["offsets"], R["offsets"])
        assert np.array_equal(L["data"], R["data"])
        assert abs(L["attrs"]["dx"] - t.dx_base / 2 ** rl) == 0
        # boxes: cubes of 8^r whole blocks tiling this level's listed blocks
        lvl_slots = [s for s in slots if a["level"][s] - bl == rl]
        covered = 0
        for (lo, hi), arr in L["blocks"]:
            side = hi[0] - lo[0] + 1
            assert side == hi[1] - lo[1] + 1 == hi[2] - lo[2] + 1 and side % E == 0
            g = side // E
            assert g & (g - 1) == 0
            covered += g ** 3
            # every cell of the box is the field on its block (independent of the writer's order)
            for s in lvl_slots:
                b = a["bbase"][s]
                if all(lo[d] <= b[d] <= hi[d] for d in range(3)):
                    o = [b[d] - lo[d] for d in range(3)]
                    blk = arr[:, o[2]:o[2] + E, o[1]:o[1] + E, o[0]:o[0] + E]
                    assert np.array_equal(blk[:3], u[:, s, 1:-1, 1:-1, 1:-1])
                    assert np.array_equal(blk[3], p[0, s, 1:-1, 1:-1, 1:-1])
        assert covered == len(lvl_slots)


def test_restart_round_trip_equals_written_fields(tmp_path):
    t = _adaptive()
    a, info, n = _export(t), t.info(), t.n
    nb = info["nblocks"]
    rng = np.random.default_rng(5)
    u = rng.standard_normal((3, nb, n, n, n))
    path = str(tmp_path / "restart.hdf5")
    t.write_chombo(path, [("u", u)])
    back = t.read_chombo(path, "u", 3, MESH_FACE, out=np.full_like(u, 7.0))
    leaf = (a["flags"] & FLAG_LEAF) != 0
    corr = (a["flags"] & FLAG_CORRECTION) != 0
    rl = a["level"] - info["base_level"]
    listed = (rl >= 0) & (~corr | (rl == 0))
    filled = leaf & listed
    assert np.array_equal(back[:, filled, 1:-1, 1:-1, 1:-1], u[:, filled, 1:-1, 1:-1, 1:-1])
    assert np.all(back[:, filled][:, :, 0] == 0)  # buffers are not in the file
    assert np.all(back[:, listed & ~leaf] == 0)  # cleared, then only leaves receive values
    assert np.all(back[:, ~listed] == 7.0)  # blocks outside the list keep their values
    fa = dict(components=["u_0", "u_1", "u_2"], dx0=t.dx_base, levels=[])
    comps, lv, _, _ = _decode(path)
    for l in range(len(lv)):
        fa["levels"].append(dict(boxes=[(tuple(lo), tuple(hi)) for (lo, hi), _ in lv[l]["blocks"]],
                                 data=lv[l]["data"]))
    exp = C.read_u(a, info["base_level"], t.dx_base, fa, "u", 3, MESH_FACE, np.full_like(u, 7.0))
    assert np.array_equal(back, exp)
    with pytest.raises(IblgfError):
        t.read_chombo(path, "w", 3, MESH_FACE)


def _positions(idx, dx, mesh, comp):
    """Physical coordinate of the value at integer index idx (3, ...) of a field."""
    off = np.full(3, 0.5)
    if mesh == MESH_FACE:
        off[comp] = 0.0
    elif mesh == MESH_EDGE:
        off[:] = 0.0
        off[comp] = 0.5
    elif mesh == MESH_VERTEX:
        off[:] = 0.0
    return [(idx[d] + off[d]) * dx for d in range(3)]


@pytest.mark.parametrize("mesh", [MESH_CELL, MESH_FACE, MESH_EDGE, MESH_VERTEX])
def test_coarser_leaves_reproduce_linear_fields(tmp_path, mesh):
    """A file one and two levels finer than the leaves that read it: linear
    fields are exact at the leaves' own positions; a quadratic along x gives
    the rule's middle-cell average."""
    E = 6
    fine = Tree((-24,) * 3, (48,) * 3, E, 1.0 / 32, [(-24, -24, -24, 48, 48, 48)])
    fine.init_refine(0)
    fa, fn = _export(fine), fine.n
    ncomp = 3 if mesh in (MESH_FACE, MESH_EDGE) else 1
    coef = np.array([[0.3, 1.7, -0.9, 0.45], [-1.1, 0.2, 0.8, -0.6], [0.7, -0.4, 0.25, 1.3]])
    g = np.indices((fn, fn, fn))[::-1] - 1  # x, y, z of every stored point relative to the block base
    vals = np.zeros((ncomp, len(fa["key"]), fn, fn, fn))
    quad = np.zeros_like(vals)
    for s in range(len(fa["key"])):
        idx = [g[d] + fa["bbase"][s][d] for d in range(3)]
        for c in range(ncomp):
            x = _positions(idx, fine.dx_base, mesh, c)
            vals[c, s] = coef[c, 0] + coef[c, 1] * x[0] + coef[c, 2] * x[1] + coef[c, 3] * x[2]
            quad[c, s] = x[0] ** 2
    path = str(tmp_path / "fine.hdf5")
    fine.write_chombo(path, [("f", vals), ("q", quad)])
    for diff in (1, 2):
        m = 2 ** diff
        lo, ext, E2 = -24 // m, 48 // m, 6 // 2 ** (diff - 1)  # 4 blocks a side: 6 cells, then 3
        coarse = Tree((lo,) * 3, (ext,) * 3, E2, fine.dx_base * m, [(lo,) * 3 + (ext,) * 3])
        coarse.init_refine(0)
        ca, cn = _export(coarse), coarse.n
        got = coarse.read_chombo(path, "f", ncomp, mesh)
        gq = coarse.read_chombo(path, "q", ncomp, mesh)
        bl = coarse.info()["base_level"]
        for s in range(len(ca["key"])):
            if not (ca["flags"][s] & FLAG_LEAF) or ca["level"][s] < bl:
                continue
            gc = np.indices((cn, cn, cn))[::-1] - 1
            idx = [gc[d] + ca["bbase"][s][d] for d in range(3)]
            inner = (slice(1, -1),) * 3
            for c in range(ncomp):
                x = _positions(idx, coarse.dx_base, mesh, c)
                want = coef[c, 0] + coef[c, 1] * x[0] + coef[c, 2] * x[1] + coef[c, 3] * x[2]
                assert np.allclose(got[c, s][inner], want[inner], rtol=0, atol=1e-13)
                # middle cells along x: 1 (node positions) or 2 of m at spacing dx_fine
                nodes_x = (mesh == MESH_FACE and c == 0) or (mesh == MESH_EDGE and c != 0) or mesh == MESH_VERTEX
                hw = 0.0 if (nodes_x or m == 1) else 0.5 * fine.dx_base
                assert np.allclose(gq[c, s][inner], (x[0] ** 2 + hw ** 2)[inner], rtol=0, atol=1e-13)
        # the restatement agrees bit for bit
        comps, lv, _, _ = _decode(path)
        fd = dict(components=comps, dx0=fine.dx_base,
                  levels=[dict(boxes=[(tuple(lo), tuple(hi)) for (lo, hi), _ in lv[0]["blocks"]], data=lv[0]["data"])])
        exp = C.read_u(ca, coarse.info()["base_level"], coarse.dx_base, fd, "f", ncomp, mesh,
                       np.zeros_like(got))
        assert np.array_equal(got, exp)


@pytest.mark.skipif(not os.path.exists(REF_HDF5), reason="the reference's ref.hdf5 is read in the build container only")
def test_reads_and_rewrites_the_reference_file(tmp_path):
    """The reference's own Chombo file: read into a tree of its cells (u as a
    face field, p as a cell field) it equals the independent box decode
    (tests/golden/ns_amr_lgf_ref.npz); written back with all eight
    components, the same decoder gives the same dense arrays."""
    ref = np.load(os.path.join(os.path.dirname(__file__), "golden", "ns_amr_lgf_ref.npz"))
    t = Tree((-12,) * 3, (24,) * 3, 6, 0.25, [(-12, -12, -12, 24, 24, 24)])
    t.init_refine(0)
    a = _export(t)
    u = t.read_chombo(REF_HDF5, "u", 3, MESH_FACE)
    p = t.read_chombo(REF_HDF5, "p", 1, MESH_CELL)
    dense_u = np.zeros((3, 24, 24, 24))
    dense_p = np.zeros((24, 24, 24))
    for s in range(len(a["key"])):
        b = a["bbase"][s] + 12
        sl = (slice(b[2], b[2] + 6), slice(b[1], b[1] + 6), slice(b[0], b[0] + 6))
        dense_u[(slice(None),) + sl] = u[:, s, 1:-1, 1:-1, 1:-1]
        dense_p[sl] = p[0, s, 1:-1, 1:-1, 1:-1]
    assert np.array_equal(dense_u, ref["u"]) and np.array_equal(dense_p, ref["p"])
    # all eight components written back
    comps, lv, _, _ = _decode(REF_HDF5)
    fields, k = [], 0
    names = [("cell_aux", 1), ("edge_aux", 3), ("u", 3), ("p", 1)]
    assert comps == ["cell_aux", "edge_aux_0", "edge_aux_1", "edge_aux_2", "u_0", "u_1", "u_2", "p"]
    for name, nc in names:
        mesh = MESH_EDGE if name == "edge_aux" else (MESH_FACE if name == "u" else MESH_CELL)
        fields.append((name, t.read_chombo(REF_HDF5, name, nc, mesh)))
    out = str(tmp_path / "again.hdf5")
    t.write_chombo(out, fields)
    comps2, lv2, _, _ = _decode(out)
    assert comps2 == comps

    def dense(levels):
        d = np.zeros((8, 24, 24, 24))
        for (lo, hi), arr in levels[0]["blocks"]:
            d[:, lo[2] + 12:hi[2] + 13, lo[1] + 12:hi[1] + 13, lo[0] + 12:hi[0] + 13] = arr
        return d

    assert np.array_equal(dense(lv), dense(lv2))
    assert int(lv2[0]["offsets"].sum()) == 8 * 24 ** 3


def test_tree_restart_file_round_trip(tmp_path):
    t = _adaptive(block_extent=8)
    a = t.export()
    path = str(tmp_path / "tree_info.bin")
    t.write_restart(path)
    raw = open(path, "rb").read()
    nb = len(a["key"])
    assert len(raw) == 9 * nb
    keys = np.frombuffer(raw[:8 * nb], dtype="<u8")
    leafs = np.frombuffer(raw[8 * nb:], dtype=np.uint8)
    # pre-order depth first: coordinate bits first, then level (a parent before its children)
    order = sorted(range(nb), key=lambda s: (int(a["key"][s]) >> 7, int(a["level"][s])))
    assert np.array_equal(keys, a["key"][order])
    assert np.array_equal(leafs, ((a["flags"][order] & FLAG_LEAF) != 0).astype(np.uint8))
    t2 = Tree.read_restart(path, (-48,) * 3, (96,) * 3, 8, 1.0 / 16)
    b = t2.export()
    assert t2.info() == t.info()
    for k in ("key", "level", "coord", "bbase", "parent", "children", "neighbor", "influence", "influence_num"):
        assert np.array_equal(a[k], b[k]), k
    m = FLAG_LEAF | FLAG_CORRECTION | 4 | 8  # leaf, correction, leaf boundary, physical: recomputed
    assert np.array_equal(a["flags"] & m, b["flags"] & m)
    with pytest.raises(IblgfError):
        Tree.read_restart(str(tmp_path / "missing.bin"), (-48,) * 3, (96,) * 3, 8, 1.0 / 16)
    open(str(tmp_path / "bad.bin"), "wb").write(raw[:-1])
    with pytest.raises(IblgfError):
        Tree.read_restart(str(tmp_path / "bad.bin"), (-48,) * 3, (96,) * 3, 8, 1.0 / 16)


def test_oracle_restart_from_the_files_steps_as_the_continued_run(tmp_path, oracle_mod):
    """The premise of a restart (the reference's ns_amr_lgf.hpp:294-307: the
    tree from its keys, then u from the restart field file): the IF-HERK step
    depends on the velocity's leaf interiors only, so the oracle's step from
    a velocity written and read back by the library equals its step from the
    velocity it kept."""
    from iblgf_amd import setups

    v = setups.VortexRings(seed=7, block_extent=8, base_blocks=4, nref=2, nrings=3, centre_range=4.0,
                           radius=(4.0, 8.0), core=(1.5, 2.5), refinement_factor=0.5)
    blobs = [((0.02, -0.03, 0.01), 0.09, (0, 0, 1), 1.0), ((-0.04, 0.02, 0.03), 0.08, (1, 0.5, 0), 0.7)]
    flow = setups.GaussianFlow(v.tree, blobs)
    ot = oracle_mod.Tree(flow.arrays, v.info, v.tree.n)
    u1, _ = ot.ifherk_step(flow.velocity(), 2e-4, 500.0, None)
    u2, p2 = ot.ifherk_step(u1, 2e-4, 500.0, None)
    tree_file, field_file = str(tmp_path / "tree_info.bin"), str(tmp_path / "restart_field.hdf5")
    v.tree.write_restart(tree_file)
    v.tree.write_chombo(field_file, [("u", np.stack(u1))])
    t2 = Tree.read_restart(tree_file, (-32,) * 3, (64,) * 3, 8, v.dx_base)
    t2.build_masks()
    assert np.array_equal(t2.export()["key"], flow.arrays["key"])
    ur = t2.read_chombo(field_file, "u", 3, MESH_FACE)
    assert not all(np.array_equal(a, b) for a, b in zip(ur, u1))  # non-leaf blocks and buffers were dropped
    u2r, p2r = ot.ifherk_step(list(ur), 2e-4, 500.0, None)
    leaves = (flow.arrays["flags"] & FLAG_LEAF) != 0
    assert all(np.array_equal(a[leaves], b[leaves]) for a, b in zip(u2r, u2))
    assert np.array_equal(p2r[leaves], p2[leaves])
