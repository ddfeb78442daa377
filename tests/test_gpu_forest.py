"""Forests of independent trees in one context (iblgf_gpu_forest_upload) and
solve sequences captured as one graph (iblgf_gpu_solve_sequence).

Every launch of a forest solve covers the blocks of all trees, but each
tree's arithmetic is its own (sibling groups, transforms and transfers are
per block or per group), so each tree's result equals its single-tree solve
bit for bit, and the oracle bounds hold per tree (relative L2 <= 1e-10 in
FP64).  A sequence equals its items' separate calls bit for bit, and its
graph replays equal the eager run.
"""
import numpy as np
import pytest

from iblgf_amd import (FLAG_LEAF, KERNEL_IF, KERNEL_LAPLACE, PREC_FP16X3_FORWARD, PREC_FP64, PREC_MIXED, PoissonSolver,
                       setups)

pytestmark = pytest.mark.gpu

SEEDS = (7, 11, 23)


def _trees():
    return [setups.VortexRings(seed=s, block_extent=8, base_blocks=4, nref=2, nrings=2 + i, centre_range=4.0,
                               radius=(4.0, 8.0), core=(1.5, 2.5), refinement_factor=0.5)
            for i, s in enumerate(SEEDS)]


_W = {}


def _workload():
    if not _W:
        ws = _trees()
        _W["ws"] = ws
        _W["srcs"] = [w.sources() for w in ws]  # [tree][3][nb_i, n, n, n]
    return _W["ws"], _W["srcs"]


def _single(w, srcs, precision, kernel, alpha):
    s = PoissonSolver(w.tree)
    s.set_precision(precision)
    fi = [s.field() for _ in range(3)]
    fo = [s.field() for _ in range(3)]
    for f, x in zip(fi, srcs):
        s.upload(f, x)
    s.apply_lgf_batch(fi, fo, kernel=kernel, alpha_base=alpha)
    return [s.download(f) for f in fo]


@pytest.mark.parametrize("precision,kernel,alpha", [(0, KERNEL_LAPLACE, 0.0), (0, KERNEL_IF, 0.7),
                                                    (PREC_MIXED, KERNEL_LAPLACE, 0.0),
                                                    (PREC_FP16X3_FORWARD, KERNEL_LAPLACE, 0.0)])
def test_forest_equals_single_tree_solves(oracle_mod, precision, kernel, alpha):
    ws, srcs = _workload()
    print(f"case: forest of {len(ws)} trees {[w.info['nblocks'] for w in ws]}, precision {precision}, "
          f"kernel {kernel}, alpha {alpha}", flush=True)
    fs = PoissonSolver([w.tree for w in ws])
    assert fs.nblocks == sum(w.info["nblocks"] for w in ws)
    fs.set_precision(precision)
    fi = [fs.field() for _ in range(3)]
    fo = [fs.field() for _ in range(3)]
    for c in range(3):
        fs.upload(fi[c], fs.to_forest([s[c] for s in srcs]))
    fs.apply_lgf_batch(fi, fo, kernel=kernel, alpha_base=alpha)
    got = [fs.from_forest(fs.download(f)) for f in fo]  # [comp][tree]
    for i, w in enumerate(ws):
        ref = _single(w, srcs[i], precision, kernel, alpha)
        for c in range(3):
            assert np.array_equal(got[c][i], ref[c]), (i, c)
    if precision == 0:
        w = ws[1]
        ot = oracle_mod.Tree(w.arrays, w.info, w.tree.n)
        ref = ot.apply_lgf(srcs[1][2]) if kernel == KERNEL_LAPLACE else ot.apply_if(srcs[1][2], alpha)
        sel = (w.arrays["flags"] & FLAG_LEAF) != 0 if kernel == KERNEL_LAPLACE else slice(None)
        d = got[2][1][sel] - ref[sel]
        rel = np.linalg.norm(d) / np.linalg.norm(ref[sel])
        print(f"tree 1 component 2 vs oracle: relL2={rel:.3e}", flush=True)
        assert rel < 1e-10


def test_solve_sequence_equals_calls_and_replays():
    ws, srcs = _workload()
    print("case: IF-HERK-like sequence (3 stages of IF + Poisson, 3 components) on the forest", flush=True)
    fs = PoissonSolver([w.tree for w in ws])
    fs.set_precision(PREC_MIXED)
    w_ = [fs.field() for _ in range(3)]
    u = [fs.field() for _ in range(3)]
    p = [fs.field() for _ in range(3)]
    for c in range(3):
        fs.upload(w_[c], fs.to_forest([s[c] for s in srcs]))
    alphas = (0.78, 0.56, 1.66)
    seq = []
    for a in alphas:
        seq.append((KERNEL_IF, a, w_, u))
        seq.append((KERNEL_LAPLACE, 0.0, u, p))
    # separate calls
    for kernel, a, src, dst in seq:
        fs.apply_lgf_batch(src, dst, kernel=kernel, alpha_base=a)
    ref_u = [fs.download(f) for f in u]
    ref_p = [fs.download(f) for f in p]
    # eager sequence
    for f in u + p:
        fs.upload(f, np.zeros(fs.shape))
    fs.solve_sequence(seq)
    assert all(np.array_equal(fs.download(f), r) for f, r in zip(u + p, ref_u + ref_p))
    # one graph per sequence, replayed with eager work in between
    fs.set_graphs(True)
    for rep in range(3):
        for f in u + p:
            fs.upload(f, np.zeros(fs.shape))
        fs.solve_sequence(seq)
        fs.sync()
        got = [fs.download(f) for f in u + p]
        for g, r in zip(got, ref_u + ref_p):
            assert np.array_equal(g, r), rep
        fs.apply_lgf_batch(w_, u)  # eager work between replays
    fs.set_graphs(False)


def test_autotuner_keeps_results_and_records_its_choice():
    """Every launch shape the autotuner tries gives the default shape's result
    bit for bit; it keeps the fastest and reports all twelve timings."""
    ws, srcs = _workload()
    print("case: autotune on the forest, mixed precision, Laplace, 3 components", flush=True)
    fs = PoissonSolver([w.tree for w in ws])
    fs.set_precision(PREC_MIXED)
    fi = [fs.field() for _ in range(3)]
    fo = [fs.field() for _ in range(3)]
    for c in range(3):
        fs.upload(fi[c], fs.to_forest([s[c] for s in srcs]))
    fs.apply_lgf_batch(fi, fo)
    ref = [fs.download(f) for f in fo]
    for mp in (0, 1, 2):
        for dg in (0, 1):
            for conc in (True, False):
                fs.set_launch_shapes(mp, dg)
                fs.set_concurrency(conc)
                fs.apply_lgf_batch(fi, fo)
                assert all(np.array_equal(fs.download(f), r) for f, r in zip(fo, ref)), (mp, dg, conc)
    report, choice = fs.autotune(fi, fo, reps=2)
    print(report, flush=True)
    lines = [l for l in report.splitlines() if l.startswith("mac_policy=")]
    assert len(lines) == 12 and report.splitlines()[-1].startswith("chosen:")
    assert f"mac_policy={choice['mac_policy']} dft_grid={choice['dft_grid']} streams={choice['streams']}" in \
        report.splitlines()[-1]
    fs.apply_lgf_batch(fi, fo)
    assert all(np.array_equal(fs.download(f), r) for f, r in zip(fo, ref))


@pytest.mark.parametrize("prec", [PREC_FP64, PREC_MIXED])
def test_work_queue_launches_repeat_and_replay(prec):
    """The persistent work-queue multiply-accumulate (mac_policy 2): its
    counters reset themselves at the end of every launch, so repeated
    Laplace and IF solves, eager and replayed from graphs, equal the default
    grid's results bit for bit."""
    ws, srcs = _workload()
    print(f"case: work-queue MAC on the forest, precision {prec}, Laplace and IF, eager and graph", flush=True)
    fs = PoissonSolver([w.tree for w in ws])
    fs.set_precision(prec)
    fi = [fs.field() for _ in range(3)]
    fo = [fs.field() for _ in range(3)]
    for c in range(3):
        fs.upload(fi[c], fs.to_forest([s[c] for s in srcs]))

    def run():
        for f in fo:  # Laplace writes leaves only: start every sequence from zeros
            fs.upload(f, np.zeros(fs.shape))
        out = []
        for kern, alpha in ((KERNEL_LAPLACE, 0.0), (KERNEL_IF, 0.4), (KERNEL_LAPLACE, 0.0)):
            fs.apply_lgf_batch(fi, fo, kernel=kern, alpha_base=alpha)
            fs.sync()
            out.append([fs.download(f) for f in fo])
        return out

    ref = run()
    fs.set_launch_shapes(2, 0)
    for graphs in (False, True):
        fs.set_graphs(graphs)
        for rep in range(3):
            got = run()
            for i, (g, r) in enumerate(zip(got, ref)):
                assert all(np.array_equal(a, b) for a, b in zip(g, r)), (graphs, rep, i)
            fs.apply_lgf(fi[0], fo[0])  # other work between sequences
    fs.set_graphs(False)


def test_fp32_fields_round_the_fp64_field_solve():
    """FP32 user fields: the solve reads them into its FP64 work arrays and
    rounds its result once on the way out, so an FP32-field solve equals the
    FP64-field solve of the same (float-valued) sources rounded to FP32, bit
    for bit; operations outside the AMR2AMR solves refuse FP32 fields."""
    from iblgf_amd import IBLGF_ERR_ARG, IblgfError

    ws, srcs = _workload()
    print("case: FP32 fields, forest, mixed precision, Laplace + IF sequence", flush=True)
    fs = PoissonSolver([w.tree for w in ws])
    fs.set_precision(PREC_MIXED)
    x = [fs.to_forest([s[c] for s in srcs]).astype(np.float32) for c in range(3)]
    w64 = [fs.field() for _ in range(3)]
    o64 = [fs.field() for _ in range(3)]
    w32 = [fs.field_f32() for _ in range(3)]
    o32 = [fs.field_f32() for _ in range(3)]
    for c in range(3):
        fs.upload(w64[c], x[c].astype(np.float64))
        fs.upload(w32[c], x[c])
    for kernel, alpha in ((KERNEL_LAPLACE, 0.0), (KERNEL_IF, 0.56)):
        fs.apply_lgf_batch(w64, o64, kernel=kernel, alpha_base=alpha)
        fs.apply_lgf_batch(w32, o32, kernel=kernel, alpha_base=alpha)
        for c in range(3):
            ref = fs.download(o64[c]).astype(np.float32)
            got = fs.download(o32[c])
            assert got.dtype == np.float32 and np.array_equal(got, ref), (kernel, c)
    with pytest.raises(IblgfError) as e:
        fs.upload(w32[0], np.zeros(fs.shape, np.float64))  # converted to float32: allowed
        from iblgf_amd import lib
        rc = lib().iblgf_gpu_buffer_exchange(fs._h, w32[0], ws[0].info["base_level"])
        if rc:
            raise IblgfError(rc, "buffer exchange")
    assert e.value.code == IBLGF_ERR_ARG
