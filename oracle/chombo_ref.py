"""Pure-Python restatement of the reference's Chombo output and restart
reader (include/iblgf/IO/chombo/chombo_impl.hpp, h5_io.hpp) over the block
table the library exports (iblgf_tree_export: slot order, level-major and
key-sorted; bbase = block base in cells of the block's refinement level).

TEST INFRASTRUCTURE ONLY: the checker the library's iblgf_chombo_write /
iblgf_chombo_read are compared against.  Imported by tests/ only.

Functions:
  blocks_list   h5_io.hpp:232-263 (blocks_list_build)
  level_groups  chombo_impl.hpp:95-199 (Chombo::init, block_consecutive)
  encode        chombo_impl.hpp:1755-2026, 2336-2598 (the values the writer
                puts in each level's boxes, offsets and data datasets)
  read_u        chombo_impl.hpp:217-626 (Chombo::read_u)
Pins (tests/test_chombo.py): the reference's own file
tests/ns_amr_lgf/configs/ref.hdf5 decoded independently
(tests/golden/make_ns_amr_lgf_golden.py) against read_u on the same cells;
exact reproduction of linear fields at every averaging factor and mesh type;
the box/group invariants of the writer.
"""
import math

import numpy as np

FLAG_LEAF, FLAG_CORRECTION = 1, 2
MESH_CELL, MESH_FACE, MESH_EDGE, MESH_VERTEX = range(4)
MAX_LEVEL, SHIFT_TO_COORD = 19, 7


def blocks_list(a, base_level, include_correction, base_correction_only):
    """Slots of blocks_list_build, per refinement level in key order (the
    depth-first order of the reference restricted to a level)."""
    out = []
    for s in range(len(a["key"])):
        rl = int(a["level"][s]) - base_level
        corr = bool(a["flags"][s] & FLAG_CORRECTION)
        if rl < 0:
            continue
        if not include_correction and corr:
            continue
        if base_correction_only and corr and rl > 0:
            continue
        out.append(s)
    return out


def _key_next(k):
    lvl = (int(k) >> 2) & 31
    return int(k) + (1 << (SHIFT_TO_COORD + 3 * (MAX_LEVEL - lvl)))


def _child_number(k):
    lvl = (int(k) >> 2) & 31
    return (int(k) >> (64 - 3 * lvl)) & 7 if lvl > 0 else 0


def _parent(k):
    lvl = (int(k) >> 2) & 31
    keep = ((1 << 64) - 1) ^ ((1 << (64 - 3 * (lvl - 1))) - 1) if lvl > 1 else 0
    return (int(k) & keep) | ((lvl - 1) << 2) | 1


def level_groups(a, base_level, slots):
    """{refinement level: (probe (base3, max3), [groups of slots])}."""
    levels = {}
    E = None
    for s in slots:
        rl = int(a["level"][s]) - base_level
        levels.setdefault(rl, []).append(s)
    out = {}
    for rl in sorted(levels):
        oc = levels[rl]
        # probe domain: the reference moves the base without its max
        base = [int(v) for v in a["bbase"][oc[0]]]
        E = E or int(a["block_extent"])
        ext = [E, E, E]
        for s in oc[1:]:
            b = [int(v) for v in a["bbase"][s]]
            for d in range(3):
                if base[d] > b[d]:
                    base[d] = b[d]
                if base[d] + ext[d] - 1 < b[d] + E - 1:
                    ext[d] = b[d] + E - 1 - base[d] + 1
        groups, track = [], 0
        for i in range(len(oc)):
            if i == track:
                groups.append([oc[i]])
                recurse, lim = 0, int(a["level"][oc[i]])
                pk = int(a["key"][oc[i]])
                while i == track and recurse < lim:
                    shift = 8 ** (recurse + 1) - 1
                    shift_final = 8 ** recurse
                    if _child_number(pk) == 0 and i + shift <= len(oc) - 1:
                        k, ok = int(a["key"][oc[i]]), True
                        for j in range(shift + 1):
                            if int(a["key"][oc[i + j]]) != k:
                                ok = False
                                break
                            k = _key_next(k)
                        if ok:
                            recurse += 1
                            pk = _parent(pk)
                        else:
                            track += shift_final
                    else:
                        track += shift_final
            elif i < track:
                groups[-1].append(oc[i])
        out[rl] = ((tuple(base), tuple(base[d] + ext[d] - 1 for d in range(3))), groups)
    return out


def encode(a, base_level, fields, include_correction=True, base_correction_only=True):
    """fields: list of (name, array [ncomp, nslots, n, n, n]).  Returns
    {level: dict(probe, boxes [(lo3, hi3)], offsets, data)} and the component
    names, as the writer lays them out."""
    E = int(a["block_extent"])
    names = []
    for name, f in fields:
        names += [name] if f.shape[0] == 1 else [f"{name}_{c}" for c in range(f.shape[0])]
    ncomp = len(names)
    slots = blocks_list(a, base_level, include_correction, base_correction_only)
    out = {}
    for rl, (probe, groups) in level_groups(a, base_level, slots).items():
        boxes, offsets, data = [], [0], []
        for grp in groups:
            offsets.append(len(grp) * E ** 3 * ncomp)
            ge = round(len(grp) ** (1 / 3))
            b0 = np.asarray(a["bbase"][grp[0]], dtype=np.int64)
            ordered = {}
            for s in grp:
                sh = (np.asarray(a["bbase"][s], dtype=np.int64) - b0) // E
                ordered[int(sh[0] + sh[1] * ge + sh[2] * ge * ge)] = s
            lo = np.min([a["bbase"][s] for s in grp], axis=0)
            hi = np.max([np.asarray(a["bbase"][s]) + E - 1 for s in grp], axis=0)
            boxes.append((tuple(int(v) for v in lo), tuple(int(v) for v in hi)))
            for _, f in fields:
                for c in range(f.shape[0]):
                    for z in range(ge):
                        for k in range(E):
                            for y in range(ge):
                                for j in range(E):
                                    for x in range(ge):
                                        s = ordered[x + y * ge + z * ge * ge]
                                        data.extend(f[c, s, k + 1, j + 1, 1:E + 1])
        out[rl] = dict(probe=probe, boxes=boxes, offsets=offsets, data=np.asarray(data))
    return out, names


def read_u(a, base_level, dx_base, file, name, nfields, mesh, data):
    """Chombo::read_u into data [nfields, nslots, n, n, n] (modified in place).
    file: dict(components [names], dx0, levels [dict(boxes [(lo3, hi3)], data)])."""
    E = int(a["block_extent"])
    n = E + 2
    slots = blocks_list(a, base_level, True, True)
    for s in slots:
        data[:, s] = 0.0
    comps = file["components"]
    cidx = None
    for i, c in enumerate(comps):
        if (nfields > 1 and c == name + "_0") or (nfields == 1 and c == name):
            cidx = i
            break
    if cidx is None:
        raise KeyError(name)
    diff_f = math.log2(dx_base / file["dx0"])
    if abs(round(diff_f) - diff_f) > 1e-10:
        raise ValueError("base dx doesn't match")
    bld = int(round(diff_f))

    def tdiv(x, f):  # C++ integer division (towards zero)
        return abs(x) // f * (1 if x >= 0 else -1)

    def scale(base, ext, lfrom, lto):
        """BlockDescriptor::level_scale (blockDescriptor.hpp:212-236) to a coarser level."""
        f = 2 ** abs(lfrom - lto)
        mx = [base[d] + ext[d] - 1 for d in range(3)]
        nb = [tdiv(base[d] + 1, f) - 1 if base[d] < 0 else base[d] // f for d in range(3)]
        nm = [tdiv(mx[d] + 1, f) - 1 if mx[d] < 0 else mx[d] // f for d in range(3)]
        return nb, [nm[d] - nb[d] + 1 for d in range(3)]

    for l, lev in enumerate(file["levels"]):
        fake = l + bld
        box_offset = 0
        for (lo, hi) in lev["boxes"]:
            fext = [hi[d] - lo[d] + 1 for d in range(3)]
            fsize = fext[0] * fext[1] * fext[2]
            for s in slots:
                if not (int(a["flags"][s]) & FLAG_LEAF):
                    continue
                rl = int(a["level"][s]) - base_level
                if rl > fake:
                    continue
                factor = 2 ** (fake - rl)
                bb = [int(v) for v in a["bbase"][s]]
                obase, oext = (list(lo), fext) if rl == fake else scale(list(lo), fext, fake, rl)
                ob = [max(bb[d], obase[d]) for d in range(3)]
                oe = [min(bb[d] + E, obase[d] + oext[d]) - ob[d] for d in range(3)]
                if min(oe) < 1:
                    continue
                up = [v * factor for v in ob]
                avg = ((factor - 1) // 2, factor // 2)
                for fi in range(nfields):
                    sep = nfields // 3 if nfields > 3 else 1
                    fv = [avg, avg, avg]
                    if mesh == MESH_FACE:
                        fv[fi // sep] = (0, 0)
                    elif mesh == MESH_EDGE:
                        fv = [(0, 0)] * 3
                        fv[fi // sep] = avg
                    elif mesh == MESH_VERTEX:
                        fv = [(0, 0)] * 3
                    total = 1.0
                    for d in range(3):
                        total *= fv[d][1] - fv[d][0] + 1
                    for sk in range(fv[2][0], fv[2][1] + 1):
                        for k in range(oe[2]):
                            for sj in range(fv[1][0], fv[1][1] + 1):
                                for j in range(oe[1]):
                                    for si in range(fv[0][0], fv[0][1] + 1):
                                        f2 = up[2] - lo[2] + k * factor + sk
                                        f1 = up[1] - lo[1] + j * factor + sj
                                        if not (0 <= f2 < fext[2] and 0 <= f1 < fext[1]):
                                            continue
                                        row = box_offset + (cidx + fi) * fsize + f2 * fext[0] * fext[1] + f1 * fext[0]
                                        for i in range(oe[0]):
                                            f0 = up[0] - lo[0] + i * factor + si
                                            if not 0 <= f0 < fext[0]:
                                                continue
                                            x = i + ob[0] - bb[0] + 1
                                            y = j + ob[1] - bb[1] + 1
                                            z = k + ob[2] - bb[2] + 1
                                            data[fi, s, z, y, x] += lev["data"][row + f0] / total
            box_offset += fsize * len(comps)
    assert n == data.shape[-1]
    return data
