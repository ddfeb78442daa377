"""Device mirror refresh after an adaptation (iblgf_gpu_tree_readapt) against
the oracle's remap (Ifherk::adapt, ifherk.hpp:1047-1074): kept blocks move by
key on the device, the children of the parents of new leaves are
interpolated (relative 1e-13; the transfer sums in another order), unlisted
fields restart at zero, and the solves and the IF-HERK step on the adapted
tree match the oracle (relative L2 1e-10) with the kernel spectra kept."""
import numpy as np
import pytest

from iblgf_amd import FLAG_CORRECTION, FLAG_LEAF, MESH_FACE, PoissonSolver, Tree, setups

pytestmark = pytest.mark.gpu


def _make_small():
    t = Tree((-48,) * 3, (96,) * 3, 8, 1.0 / 48, [(-16, -16, -16, 32, 32, 32)])

    def cond(b, ext, rl, diff):
        c = [(b[d] + ext / 2) / 2 ** rl for d in range(3)]
        return np.linalg.norm(c) < 10 / (1 + rl)

    t.init_refine(2, cond)
    t.build_masks()
    return t


def _rel(a, b):
    return np.sqrt(sum(np.sum((x - y) ** 2) for x, y in zip(a, b))) / np.sqrt(sum(np.sum(y ** 2) for y in b))


def test_readapt_matches_oracle_remap_and_solves(oracle_mod):
    print("case: adapt a 507-block tree (refine 2 leaves, 2:1 cascade), remap u, solve and step", flush=True)
    t = _make_small()
    a = t.export()
    info, n = t.info(), t.n
    rng = np.random.default_rng(11)
    u = [rng.standard_normal((info["nblocks"], n, n, n)) for _ in range(3)]
    s = PoissonSolver(t)
    fu = [s.field() for _ in range(3)]
    fw = s.field()
    for f, x in zip(fu, u):
        s.upload(f, x)
    s.upload(fw, u[0])
    src0 = setups.GaussianFlow(t, [((0.02, -0.03, 0.01), 0.09, (0, 0, 1), 1.0)]).velocity()[0]
    s.upload(fw, src0)
    fo = s.field()
    s.apply_lgf(fw, fo)  # kernel spectra built on the old tree
    bl = info["base_level"]
    fin = [int(k) for k, f, l in zip(a["key"], a["flags"], a["level"])
           if (f & FLAG_LEAF) and not (f & FLAG_CORRECTION) and l == bl + 2]
    import ctypes

    from iblgf_amd import lib

    def coord(k):
        c = (ctypes.c_int * 3)()
        lib().iblgf_key_coord(ctypes.c_uint64(k), c)
        return np.array(list(c))

    fin.sort(key=lambda k: np.linalg.norm(coord(k) - 32.0))
    t.adapt({k: 1 for k in fin[:2]})
    t.build_masks()
    a2, info2 = t.export(), t.info()
    ot = oracle_mod.Tree(a2, info2, n)
    ref, parents = oracle_mod.remap(a, ot, u, [1, 1, 1], [0, 1, 2])
    npar = s.readapt(t, fu, MESH_FACE, [0, 1, 2])
    assert npar == len(parents) and s.shape[0] == info2["nblocks"]
    got = [s.download(f) for f in fu]
    err = _rel(got, ref)
    old = {int(k): i for i, k in enumerate(a["key"])}
    # kept blocks that no interpolation wrote: moved bit for bit
    kept = [i for i, k in enumerate(a2["key"]) if int(k) in old and int(a2["parent"][i]) not in parents]
    inner = (kept,) + (slice(1, -1),) * 3
    same = all(np.array_equal(g[inner], r[inner]) for g, r in zip(got, ref))
    print(f"{info['nblocks']} -> {info2['nblocks']} blocks, {npar} parents interpolated, remap relL2 {err:.2e}, "
          f"kept interiors bitwise {same}", flush=True)
    assert err < 1e-13 and same
    assert not s.download(fw).any()
    # solves on the adapted tree
    src = setups.GaussianFlow(t, [((0.02, -0.03, 0.01), 0.09, (0, 0, 1), 1.0)]).velocity()[0]
    s.upload(fw, src)
    s.apply_lgf(fw, fo)
    leaves = (a2["flags"] & FLAG_LEAF) != 0
    ref_phi = ot.apply_lgf(src)
    e_phi = _rel([s.download(fo)[leaves]], [ref_phi[leaves]])
    print(f"Laplace solve on the adapted tree vs oracle: relL2 {e_phi:.2e}", flush=True)
    assert e_phi < 1e-10
    # the IF-HERK step on the adapted tree
    flow = setups.GaussianFlow(t, [((0.02, -0.03, 0.01), 0.09, (0, 0, 1), 1.0),
                                   ((-0.04, 0.02, 0.03), 0.08, (1, 0.5, 0), 0.7)])
    v0 = flow.velocity()
    for f, x in zip(fu, v0):
        s.upload(f, x)
    s.ifherk_step(fu, 2e-4, 500.0)
    ref_v, _ = ot.ifherk_step(v0, 2e-4, 500.0)
    e_v = _rel([s.download(f)[leaves] for f in fu], [x[leaves] for x in ref_v])
    print(f"IF-HERK step on the adapted tree vs oracle: relL2 {e_v:.2e}", flush=True)
    assert e_v < 1e-10
