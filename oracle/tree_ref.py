"""Pure-Python restatement of the octree indexing of the LGF path.

TEST INFRASTRUCTURE ONLY (the checker).  Integer arithmetic on Python ints,
written from the reference's definitions, to compare bit-exactly with the
host octree of the product:

  Morton key            include/iblgf/domain/octree/key.hpp:57-75, 119-132, 254-293, 447-469
  Bitmasks              include/iblgf/domain/octree/bitmasks.hpp:56-77, 127-173
  neighbour keys        key.hpp:214-223 (x fastest, offsets -1..1)
  influence keys        key.hpp:226-250
  list construction     include/iblgf/domain/octree/tree.hpp:571-603
  AMR2AMR masks         include/iblgf/fmm/fmm.hpp:392-680
"""
MAX_LEVEL = 19
FLAG_BITS = 2
LEVEL_BITS = 5
SHIFT = FLAG_BITS + LEVEL_BITS
M64 = (1 << 64) - 1


def clog2(n):
    return 1 if n < 2 else 1 + clog2(n >> 1)


def spread(w):
    """Interleave: bit i of the 21 low bits of w goes to bit 3i."""
    w &= (1 << 21) - 1
    out = 0
    for i in range(21):
        if (w >> i) & 1:
            out |= 1 << (3 * i)
    return out


def compact(w):
    out = 0
    for i in range(21):
        if (w >> (3 * i)) & 1:
            out |= 1 << i
    return out


def make(x, y, z, level):
    """Key(coordinate, level): coordinates taken mod 2^21 then shifted so
    bits above the level fall off the 64-bit word."""
    idx = spread(x & M64) | (spread(y & M64) << 1) | (spread(z & M64) << 2)
    idx = (idx << SHIFT) & M64
    idx = (idx << ((MAX_LEVEL - level) * 3)) & M64
    return idx | (level << FLAG_BITS) | 1


def level(k):
    return (k >> FLAG_BITS) & ((1 << LEVEL_BITS) - 1)


def coord(k):
    L = level(k)
    sh = SHIFT + (MAX_LEVEL - L) * 3
    return tuple(compact(k >> (sh + d)) for d in range(3))


def coord_mask(l):
    return 0 if l <= 0 else (M64 << (64 - 3 * l)) & M64


def parent(k):
    L = level(k)
    if L == 0:
        return k
    return ((coord_mask(L - 1) & k) | ((L - 1) << FLAG_BITS)) + 1


def child(k, i):
    L = level(k)
    if L == MAX_LEVEL:
        return k
    return ((coord_mask(L) & k) | (i << ((MAX_LEVEL - (L + 1)) * 3 + SHIFT)) | ((L + 1) << FLAG_BITS)) + 1


def child_number(k, l):
    if l == 0:
        return -1
    return (k >> (SHIFT + 3 * (MAX_LEVEL - l))) & 7


def end_key(l):
    return ((((l << 2) | 1) | (M64 & ~((1 << SHIFT) - 1))) + 1) & M64


def is_end(k):
    return k >= end_key(level(k))


def minimum_level(m):
    L = 0
    while m >= (1 << L):
        L += 1
    return L


def neighbor_keys(k):
    c = coord(k)
    L = level(k)
    return [make(c[0] + dx, c[1] + dy, c[2] + dz, L) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def influence_keys(k):
    if level(k) == 0:
        return []
    c = coord(k)
    out = []
    for pk in neighbor_keys(parent(k)):
        if is_end(pk):
            continue
        for ci in range(8):
            ck = child(pk, ci)
            cc = coord(ck)
            if abs(cc[0] - c[0]) > 1 or abs(cc[1] - c[1]) > 1 or abs(cc[2] - c[2]) > 1:
                out.append(ck)
    return out


def lists(keys):
    """Neighbour (27 slots) and influence lists over the blocks with data, by slot."""
    slot = {int(k): s for s, k in enumerate(keys)}
    nbr, infl = [], []
    for k in keys:
        k = int(k)
        nbr.append([slot.get(nk, -1) for nk in neighbor_keys(k)])
        if level(k) == 0:
            infl.append([])
        else:
            infl.append([slot[ik] for ik in influence_keys(k) if ik in slot])
    return nbr, infl


def amr2amr_masks(a, base_level, depth, subtract_non_leaf=False):
    """FmmMaskBuilder::fmm_lgf_mask_build on an exported block table.
    Returns (mask_src, mask_tgt) as lists of Python ints (bit = mask index)."""
    nb = len(a["key"])
    lvl = [int(x) for x in a["level"]]
    leaf = [bool(f & 1) for f in a["flags"]]
    corr = [bool(f & 2) for f in a["flags"]]
    children = [[int(c) for c in row] for row in a["children"]]
    parent_ = [int(p) for p in a["parent"]]
    infl = [[int(x) for x in a["influence"][s][: a["influence_num"][s]]] for s in range(nb)]
    by_level = {}
    for s in range(nb):
        by_level.setdefault(lvl[s], []).append(s)
    S = [0] * nb
    T = [0] * nb

    def get(m, s, i):
        return (m[s] >> i) & 1

    def put(m, s, i, v):
        if v:
            m[s] |= 1 << i
        else:
            m[s] &= ~(1 << i)

    for L in range(base_level, depth):
        for nl in (0, 1):
            idx = 2 * (L - base_level) + nl + 15
            lv = by_level.get(L, [])
            for s in lv:
                put(S, s, idx, 0)
                put(T, s, idx, 0)
            for s in lv:
                if nl:
                    cpar = any(c >= 0 and corr[c] for c in children[s])
                    if subtract_non_leaf:
                        if leaf[s] or corr[s]:
                            put(S, s, idx, 0)
                            put(T, s, idx, cpar)
                        else:
                            put(S, s, idx, 1)
                            put(T, s, idx, 1)
                    else:
                        if leaf[s] or corr[s]:
                            put(S, s, idx, 1)
                            put(T, s, idx, leaf[s])
                        else:
                            put(S, s, idx, 0)
                            put(T, s, idx, 1)
                else:
                    if subtract_non_leaf:
                        put(S, s, idx, 1)
                        put(T, s, idx, 1)
                    elif leaf[s] or corr[s]:
                        put(S, s, idx, 0)
                        put(T, s, idx, leaf[s])
                    else:
                        put(S, s, idx, 1)
                        put(T, s, idx, 0)
            # upward pass
            for l in range(L - 1, -1, -1):
                for s in by_level.get(l, []):
                    put(S, s, idx, any(c >= 0 and get(S, c, idx) for c in children[s]))
                for s in by_level.get(l, []):
                    put(T, s, idx, any(c >= 0 and get(T, c, idx) for c in children[s]))
            # clean sources/targets without partners in the influence list
            for l in range(0, L):
                for s in by_level.get(l, []):
                    p = parent_[s]
                    if get(S, s, idx):
                        if p >= 0 and get(S, p, idx):
                            continue
                        if not any(get(T, x, idx) for x in infl[s]):
                            put(S, s, idx, 0)
                    if get(T, s, idx):
                        if p >= 0 and get(T, p, idx):
                            continue
                        if not any(get(S, x, idx) for x in infl[s]):
                            put(T, s, idx, 0)
    return S, T
