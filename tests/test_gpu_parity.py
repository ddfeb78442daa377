"""Parity of the HIP LGF path against the CPU FP64 oracle (needs an MI355X).

Tolerance: both sides compute in FP64; the device sums in a different order
(spectral MAC per sibling group, DFT legs as dense sums), so fields agree to
relative L2 <= 1e-10 and relative max error <= 1e-9 against the oracle.  The
single-level case is additionally checked against the brute-force direct sum.
"""
import numpy as np
import pytest

from iblgf_amd import FLAG_LEAF, KERNEL_IF, PoissonSolver, setups

pytestmark = pytest.mark.gpu

REL_L2 = 1e-10
REL_MAX = 1e-9


def _leaf_mask(arrays):
    return (arrays["flags"] & FLAG_LEAF) != 0


def _rel(a, b):
    d = a - b
    return np.linalg.norm(d) / np.linalg.norm(b), np.max(np.abs(d)) / np.max(np.abs(b))


def _small_rings(**kw):
    base = dict(seed=7, block_extent=8, base_blocks=4, nref=2, nrings=3, centre_range=4.0, radius=(4.0, 8.0),
                core=(1.5, 2.5), refinement_factor=0.5)
    base.update(kw)
    return setups.VortexRings(**base)


def test_uniform_gaussians_matches_oracle_and_direct_sum(oracle_mod):
    print("case: uniform 4^3 blocks of 8^3, seed 1000", flush=True)
    g = setups.UniformGaussians(seed=1000)
    src = g.source()
    solver = PoissonSolver(g.tree)
    fs, fd = solver.field(), solver.field()
    solver.upload(fs, src)
    solver.apply_lgf(fs, fd)
    got = solver.download(fd)
    ot = oracle_mod.Tree(g.arrays, g.info, g.tree.n)
    ref = ot.apply_lgf(src)
    leaves = _leaf_mask(g.arrays)
    l2, mx = _rel(got[leaves], ref[leaves])
    print(f"relL2={l2:.3e} relmax={mx:.3e}")
    assert l2 < REL_L2 and mx < REL_MAX
    # independent pin: brute-force direct LGF sum on two blocks
    a, n = g.arrays, g.tree.n
    sx, q = [], []
    for s in np.nonzero(leaves)[0]:
        b = a["bbase"][s]
        kk, jj, ii = np.meshgrid(np.arange(1, n - 1), np.arange(1, n - 1), np.arange(1, n - 1), indexing="ij")
        sx.append(np.stack([b[0] + ii - 1, b[1] + jj - 1, b[2] + kk - 1], -1).reshape(-1, 3))
        q.append(src[s, 1:-1, 1:-1, 1:-1].reshape(-1))
    sx, q = np.concatenate(sx), np.concatenate(q)
    for s in np.nonzero(leaves)[0][[0, 37]]:
        b = a["bbase"][s]
        kk, jj, ii = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
        tx = np.stack([b[0] + ii - 1, b[1] + jj - 1, b[2] + kk - 1], -1).reshape(-1, 3)
        direct = oracle_mod.direct_sum(sx, q, tx, g.info["dx_base"] ** 2)
        assert np.max(np.abs(got[s].reshape(-1) - direct)) < 1e-9 * np.max(np.abs(direct))


@pytest.mark.parametrize("component", [0, 2])
def test_adaptive_rings_matches_oracle(oracle_mod, component):
    print(f"case: small adaptive rings, component {component}", flush=True)
    v = _small_rings()
    src = v.source(component)
    solver = PoissonSolver(v.tree)
    fs, fd = solver.field(), solver.field()
    solver.upload(fs, src)
    solver.apply_lgf(fs, fd)
    got = solver.download(fd)
    ref = oracle_mod.Tree(v.arrays, v.info, v.tree.n).apply_lgf(src)
    leaves = _leaf_mask(v.arrays)
    l2, mx = _rel(got[leaves], ref[leaves])
    print(f"relL2={l2:.3e} relmax={mx:.3e}")
    assert l2 < REL_L2 and mx < REL_MAX


def test_adaptive_n8_matches_oracle(oracle_mod):
    print("case: adaptive rings, block extent 6 (n=8)", flush=True)
    v = _small_rings(block_extent=6, base_blocks=6, nrings=4, centre_range=6.0)
    src = v.source(1)
    solver = PoissonSolver(v.tree)
    fs, fd = solver.field(), solver.field()
    solver.upload(fs, src)
    solver.apply_lgf(fs, fd)
    got = solver.download(fd)
    ref = oracle_mod.Tree(v.arrays, v.info, v.tree.n).apply_lgf(src)
    leaves = _leaf_mask(v.arrays)
    l2, mx = _rel(got[leaves], ref[leaves])
    print(f"relL2={l2:.3e} relmax={mx:.3e}")
    assert l2 < REL_L2 and mx < REL_MAX


@pytest.mark.parametrize("extent", [10, 14, 16])
def test_adaptive_block_sizes_match_oracle(oracle_mod, extent):
    """The other supported block sizes (n = 12, 16, 18; 18 is the adaptive
    vortex-ring workload's): FP64 against the oracle, and the batched mixed
    path within relative L2 1e-6 of it on every component."""
    print(f"case: adaptive rings, block extent {extent} (n={extent + 2})", flush=True)
    v = _small_rings(block_extent=extent)
    srcs = [v.source(c) for c in range(3)]
    solver = PoissonSolver(v.tree)
    fin = [solver.field() for _ in srcs]
    fout = [solver.field() for _ in srcs]
    for f, x in zip(fin, srcs):
        solver.upload(f, x)
    ot = oracle_mod.Tree(v.arrays, v.info, v.tree.n)
    ref = ot.apply_lgf(srcs[1])
    leaves = _leaf_mask(v.arrays)
    solver.apply_lgf(fin[1], fout[1])
    l2, mx = _rel(solver.download(fout[1])[leaves], ref[leaves])
    print(f"fp64 relL2={l2:.3e} relmax={mx:.3e}", flush=True)
    assert l2 < REL_L2 and mx < REL_MAX
    solver.set_precision(1)
    solver.apply_lgf_batch(fin, fout)
    for c in range(3):
        refc = ref if c == 1 else ot.apply_lgf(srcs[c])
        l2m, mxm = _rel(solver.download(fout[c])[leaves], refc[leaves])
        print(f"mixed batched component {c}: relL2={l2m:.3e} relmax={mxm:.3e}", flush=True)
        assert l2m < 1e-6


@pytest.mark.parametrize("alpha,extent", [(0.1, 8), (1.7, 8), (0.1, 16), (1.7, 16)])
def test_integrating_factor_matches_oracle(oracle_mod, alpha, extent):
    print(f"case: IF alpha={alpha}, block extent {extent}", flush=True)
    v = _small_rings(block_extent=extent)
    src = v.source(0)
    solver = PoissonSolver(v.tree)
    fs, fd = solver.field(), solver.field()
    solver.upload(fs, src)
    solver.apply_lgf_IF(fs, fd, alpha)
    got = solver.download(fd)
    ref = oracle_mod.Tree(v.arrays, v.info, v.tree.n).apply_if(src, alpha)
    l2, mx = _rel(got, ref)
    print(f"relL2={l2:.3e} relmax={mx:.3e}")
    assert l2 < REL_L2 and mx < REL_MAX


@pytest.mark.parametrize("case", ["uniform", "rings"])
def test_mixed_precision_error(oracle_mod, case):
    """Mixed path (source and kernel spectra stored FP32, MAC products in FP32
    accumulated in FP64; DFT arithmetic, fields and level sums FP64) against
    the FP64 oracle: relative L2 <= 1e-6 per field component."""
    from iblgf_amd import PREC_MIXED

    print(f"case: mixed precision, {case}", flush=True)
    if case == "uniform":
        g = setups.UniformGaussians(seed=1000)
        srcs, tree, arrays, info = [g.source()], g.tree, g.arrays, g.info
    else:
        v = _small_rings()
        srcs, tree, arrays, info = [v.source(c) for c in range(3)], v.tree, v.arrays, v.info
    solver = PoissonSolver(tree)
    solver.set_precision(PREC_MIXED)
    fs, fd = solver.field(), solver.field()
    ot = oracle_mod.Tree(arrays, info, tree.n)
    leaves = _leaf_mask(arrays)
    for c, src in enumerate(srcs):
        solver.upload(fs, src)
        solver.apply_lgf(fs, fd)
        got = solver.download(fd)
        ref = ot.apply_lgf(src)
        l2, mx = _rel(got[leaves], ref[leaves])
        print(f"component {c}: relL2={l2:.3e} relmax={mx:.3e}")
        assert l2 < 1e-6


@pytest.mark.parametrize("kernel,prec", [(0, 0), (1, 0), (0, 1)])
def test_component_batch_equals_single_solves(oracle_mod, kernel, prec):
    """Three components in one batched solve: every launch covers all three,
    and each component's arithmetic is the single-component solve's in the
    same order, so the results are bitwise equal; component 1 is also checked
    against the oracle."""
    from iblgf_amd import KERNEL_IF

    print(f"case: component batch kernel={kernel} precision={prec}", flush=True)
    v = _small_rings()
    srcs = [v.source(c) for c in range(3)]
    solver = PoissonSolver(v.tree)
    solver.set_precision(prec)
    fs = [solver.field() for _ in range(3)]
    fb = [solver.field() for _ in range(3)]
    f1 = solver.field()
    for f, a in zip(fs, srcs):
        solver.upload(f, a)
    alpha = 0.6
    solver.apply_lgf_batch(fs, fb, kernel, alpha)
    batched = [solver.download(f) for f in fb]
    for c in range(3):
        solver.upload(f1, np.zeros(solver.shape))
        if kernel == KERNEL_IF:
            solver.apply_lgf_IF(fs[c], f1, alpha)
        else:
            solver.apply_lgf(fs[c], f1)
        assert np.array_equal(solver.download(f1), batched[c]), f"component {c} differs from its single solve"
    ot = oracle_mod.Tree(v.arrays, v.info, v.tree.n)
    if kernel == KERNEL_IF:
        ref, sel = ot.apply_if(srcs[1], alpha), slice(None)
    else:
        ref, sel = ot.apply_lgf(srcs[1]), _leaf_mask(v.arrays)
    l2, mx = _rel(batched[1][sel], ref[sel])
    print(f"component 1 vs oracle relL2={l2:.3e} relmax={mx:.3e}")
    assert l2 < (REL_L2 if prec == 0 else 1e-6)
    with pytest.raises(RuntimeError):
        solver.apply_lgf_batch([fs[0], fs[1]], [fb[0], fs[1]])


@pytest.mark.parametrize("kind", ["laplace", "if", "laplace_batch"])
def test_solve_independent_of_earlier_solves(kind):
    """The solver clears only the work blocks that a solve accumulates into
    before writing; the rest is overwritten whole.  A solve that follows a
    solve of random data in every block (halos and non-leaf blocks included)
    must give bitwise the result of a fresh solver."""
    print(f"case: earlier solve of noise, then rings, {kind}", flush=True)
    v = _small_rings()
    srcs = [v.source(0), v.source(1)]
    rng = np.random.default_rng(21)

    def run(solver, inputs):
        fin = [solver.field() for _ in inputs]
        fout = [solver.field() for _ in inputs]
        for f, x in zip(fin, inputs):
            solver.upload(f, x)
        if kind == "laplace":
            for a, b in zip(fin, fout):
                solver.apply_lgf(a, b)
        elif kind == "if":
            for a, b in zip(fin, fout):
                solver.apply_lgf_IF(a, b, 0.9)
        else:
            solver.apply_lgf_batch(fin, fout)
        return [solver.download(f) for f in fout]

    fresh = run(PoissonSolver(v.tree), srcs)
    used = PoissonSolver(v.tree)
    run(used, [1e3 * rng.standard_normal(srcs[0].shape) for _ in srcs])
    again = run(used, srcs)
    for a, b in zip(again, fresh):
        assert np.array_equal(a, b)


def test_mac_plan_statistics_are_consistent():
    """iblgf_gpu_mac_counts partitions the multiply-accumulate entries: full
    entries plus partial ones by target count add up to all entries, their
    pairs add up to the solve's pair count, and no entry carries more than 8."""
    v = _small_rings()
    solver = PoissonSolver(v.tree)
    fs, fd = solver.field(), solver.field()
    solver.upload(fs, v.source(0))
    solver.apply_lgf(fs, fd)
    solver.sync()
    m, c = solver.mac_counts(), solver.counts()
    print(f"mac plan: {m}", flush=True)
    masked = m["masked_by_targets"]
    assert m["pairs"] == c["n_pair"] > 0
    assert m["full_entries"] + sum(masked) == m["entries"]
    assert m["full_pairs"] + sum((t + 1) * n for t, n in enumerate(masked)) == m["pairs"]
    assert m["full_entries"] <= m["full_pairs"] <= 8 * m["full_entries"]
    assert m["entries"] <= m["pairs"] <= 8 * m["entries"]


@pytest.mark.parametrize("kind,prec", [("laplace_batch", 0), ("laplace_batch", 1), ("if", 0)])
def test_two_streams_match_one_stream(kind, prec):
    """Each FMM pass's coarse chain on a second stream gives bit for bit the
    result of one stream (the chains touch disjoint blocks)."""
    print(f"case: concurrency on/off, {kind}, precision {prec}", flush=True)
    v = _small_rings()
    srcs = [v.source(c) for c in range(3)]
    solver = PoissonSolver(v.tree)
    solver.set_precision(prec)
    fin = [solver.field() for _ in srcs]
    fout = [solver.field() for _ in srcs]
    for f, x in zip(fin, srcs):
        solver.upload(f, x)
    out = {}
    for on in (False, True, False, True):
        solver.set_concurrency(on)
        if kind == "if":
            for a, b in zip(fin, fout):
                solver.apply_lgf_IF(a, b, 0.9)
        else:
            solver.apply_lgf_batch(fin, fout)
        got = [solver.download(f) for f in fout]
        if on in out:
            assert all(np.array_equal(a, b) for a, b in zip(got, out[on]))
        out[on] = got
    for a, b in zip(out[True], out[False]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("prec,kernel", [(0, "laplace"), (1, "laplace"), (0, "if")])
def test_graph_replay_matches_eager(prec, kernel):
    """Batched solves replayed from a captured HIP graph equal the eager
    solves bit for bit: one capture, then three replays of the same graph,
    with eager solves of other fields, synchronisations and new source values
    between them, each compared with an eager solve of the same inputs (kernel
    timing on makes a call run eagerly without dropping the graph)."""
    print(f"case: graph replay, 3 components, precision {prec}, kernel {kernel}", flush=True)
    kw = dict(kernel=KERNEL_IF, alpha_base=0.8) if kernel == "if" else {}
    v = _small_rings()
    rng = np.random.default_rng(31)
    solver = PoissonSolver(v.tree)
    solver.set_precision(prec)
    solver.set_graphs(True)
    fs = [solver.field() for _ in range(3)]
    fg = [solver.field() for _ in range(3)]
    fe = [solver.field() for _ in range(3)]
    scratch_in, scratch_out = solver.field(), solver.field()
    base = [v.source(c) for c in range(3)]
    for rep in range(4):
        srcs = base if rep == 0 else [b * (1.0 + 0.1 * rep) + 0.01 * rng.standard_normal(solver.shape) for b in base]
        for f, a in zip(fs, srcs):
            solver.upload(f, a)
        solver.timing(True)                          # eager reference of the same inputs
        solver.apply_lgf_batch(fs, fe, **kw)
        solver.timing(False)
        eager = [solver.download(f) for f in fe]
        solver.apply_lgf_batch(fs, fg, **kw)         # rep 0 captures, later reps replay
        solver.sync()
        solver.timing(True)                          # eager work between replays
        solver.upload(scratch_in, srcs[0])
        solver.apply_lgf(scratch_in, scratch_out)
        solver.timing(False)
        solver.sync()
        replay = [solver.download(f) for f in fg]
        for c in range(3):
            assert np.array_equal(replay[c], eager[c]), f"replay {rep} component {c} differs from eager"
    solver.set_graphs(False)


def test_full_size_vortex_rings_properties():
    """Size-independent properties at the full adaptive vortex-ring size
    (4290 blocks): linearity, repeatability and agreement with the oracle's
    FMM on a sample is covered by the smaller trees above."""
    print("case: full vortex-ring tree, linearity + repeatability", flush=True)
    v = setups.VortexRings()
    s0, s1 = v.source(0), v.source(1)
    solver = PoissonSolver(v.tree)
    f0, f1, f2, d0, d1, d2, d3 = (solver.field() for _ in range(7))
    solver.upload(f0, s0)
    solver.upload(f1, s1)
    solver.upload(f2, 2.0 * s0 - 3.0 * s1)
    for fsrc, fdst in ((f0, d0), (f1, d1), (f2, d2), (f0, d3)):
        solver.apply_lgf(fsrc, fdst)
    r0, r1, r2, r3 = (solver.download(d) for d in (d0, d1, d2, d3))
    leaves = _leaf_mask(v.arrays)
    assert np.array_equal(r0[leaves], r3[leaves]), "repeated solve is not bitwise identical"
    lin = 2.0 * r0 - 3.0 * r1
    l2, _ = _rel(r2[leaves], lin[leaves])
    print(f"linearity relL2={l2:.3e}")
    assert l2 < 1e-12


@pytest.mark.parametrize("extent", [8, 16])
def test_stream_type_matches_oracle(oracle_mod, extent):
    """apply_lgf with fmm_type STREAM (the stream-function solve of
    Ifherk::pad_velocity, ifherk.hpp:1555) through the typed entry point:
    base-level pass only, written on the base-level leaves; relative L2 <= 1e-10
    against the oracle, whose STREAM solve equals the direct LGF sum over the
    non-correction base-level cells."""
    from iblgf_amd import FMM_STREAM

    v = setups.VortexRings(seed=7, block_extent=extent, base_blocks=4, nref=2, nrings=3, centre_range=4.0,
                           radius=(4.0, 8.0), core=(1.5, 2.5), refinement_factor=0.5)
    print(f"case: STREAM solve, block extent {extent}, {v.info['nblocks']} blocks", flush=True)
    rng = np.random.default_rng(17)
    n = v.tree.n
    src = rng.standard_normal((v.info["nblocks"], n, n, n))
    solver = PoissonSolver(v.tree)
    fs, fd = solver.field(), solver.field()
    solver.upload(fs, src)
    solver.upload(fd, np.full_like(src, 7.0))
    solver.apply_lgf_typed(fs, fd, fmm_type=FMM_STREAM)
    got = solver.download(fd)
    ref = oracle_mod.Tree(v.arrays, v.info, n).apply_lgf_masked(src, 0)
    a = v.arrays
    sel = (a["level"] == v.info["base_level"]) & ((a["flags"] & FLAG_LEAF) != 0)
    rel = np.linalg.norm(got[sel] - ref[sel]) / np.linalg.norm(ref[sel])
    print(f"STREAM vs oracle relL2={rel:.3e}", flush=True)
    assert rel < 1e-10
    assert np.all(got[~sel] == 7.0)  # nothing else is written


@pytest.mark.parametrize("kernel", ["laplace", "if"])
def test_face_and_edge_fields_match_oracle(oracle_mod, kernel):
    """Face and edge fields (iblgf_gpu_apply_lgf_mesh): the level transfers
    take each axis's relative position from the mesh type and the component's
    field index (cell_center_nli_intrp.hpp:599-618).  Every component matches
    the oracle's solve of that mesh type and field index (relative L2 <=
    1e-10); the cell mesh through the same entry equals the plain batched
    solve bit for bit; a sequence item with a mesh type equals the call."""
    from iblgf_amd import KERNEL_IF, KERNEL_LAPLACE, MESH_CELL, MESH_EDGE, MESH_FACE, lib

    kern, alpha = (KERNEL_LAPLACE, 0.0) if kernel == "laplace" else (KERNEL_IF, 0.7)
    print(f"case: face/edge fields, small adaptive rings, kernel {kernel}", flush=True)
    v = _small_rings()
    srcs = [v.source(c) for c in range(3)]
    solver = PoissonSolver(v.tree)
    fs = [solver.field() for _ in range(3)]
    fd = [solver.field() for _ in range(3)]
    fc = [solver.field() for _ in range(3)]
    for f, a in zip(fs, srcs):
        solver.upload(f, a)
    ot = oracle_mod.Tree(v.arrays, v.info, v.tree.n)
    sel = _leaf_mask(v.arrays) if kern == KERNEL_LAPLACE else slice(None)
    results = {}
    for mesh in (MESH_FACE, MESH_EDGE):
        solver.apply_lgf_batch(fs, fd, kern, alpha, mesh=mesh)
        got = [solver.download(f) for f in fd]
        results[mesh] = got
        for c in range(3):
            ref = ot.apply_lgf_mesh(srcs[c], mesh, c) if kern == KERNEL_LAPLACE else \
                ot.apply_if_mesh(srcs[c], alpha, mesh, c)
            l2, mx = _rel(got[c][sel], ref[sel])
            print(f"mesh {mesh} component {c}: relL2={l2:.3e} relmax={mx:.3e}", flush=True)
            assert l2 < REL_L2 and mx < REL_MAX, (mesh, c)
    # the two mesh types differ (the transfers moved), and neither is the cell solve
    solver.apply_lgf_batch(fs, fc, kern, alpha)
    cell = [solver.download(f) for f in fc]
    assert not np.array_equal(results[MESH_FACE][0], cell[0])
    assert not np.array_equal(results[MESH_FACE][0], results[MESH_EDGE][0])
    # the cell mesh through the mesh entry is the batched solve
    import ctypes
    a = (ctypes.c_int * 3)(*fs)
    b = (ctypes.c_int * 3)(*fd)
    assert lib().iblgf_gpu_apply_lgf_mesh(solver._h, a, b, 3, kern, alpha, MESH_CELL) == 0
    assert all(np.array_equal(solver.download(f), r) for f, r in zip(fd, cell))
    # sequence items carry their mesh type
    solver.solve_sequence([(kern, alpha, fs, fd, MESH_EDGE), (kern, alpha, fs, fc, MESH_FACE)])
    assert all(np.array_equal(solver.download(f), r) for f, r in zip(fd, results[MESH_EDGE]))
    assert all(np.array_equal(solver.download(f), r) for f, r in zip(fc, results[MESH_FACE]))
    assert lib().iblgf_gpu_apply_lgf_mesh(solver._h, a, b, 3, kern, alpha, 3) != 0
